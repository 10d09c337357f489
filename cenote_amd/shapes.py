"""Procedural test solids (crates/cenote/src/scene/shapes.rs) plus the UV and
geodesic spheres the benchmark scenes are built from."""

from __future__ import annotations

import numpy as np

from .scene import MeshData


def _icosahedron():
    phi = np.float32((1.0 + np.sqrt(np.float32(5.0))) / 2.0)
    p = np.array([
        [-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0],
        [0, -1, phi], [0, 1, phi], [0, -1, -phi], [0, 1, -phi],
        [phi, 0, -1], [phi, 0, 1], [-phi, 0, -1], [-phi, 0, 1],
    ], dtype=np.float32)
    p = p / np.linalg.norm(p, axis=1, keepdims=True)
    t = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.uint32)
    return p.astype(np.float32), t


def icosphere(subdivisions: int) -> MeshData:
    """scene/shapes.rs:21-43: 20*4^n triangles, shared midpoints, unit radius."""
    positions, tris = _icosahedron()
    positions = [tuple(v) for v in positions]
    for _ in range(subdivisions):
        cache = {}
        faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                pa = np.array(positions[a], dtype=np.float32)
                pb = np.array(positions[b], dtype=np.float32)
                m = (pa + pb) * np.float32(0.5)
                m = m / np.float32(np.linalg.norm(m))
                cache[key] = len(positions)
                positions.append(tuple(m))
            return cache[key]

        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            faces += [[a, ab, ca], [ab, b, bc], [ca, bc, c], [ab, bc, ca]]
        tris = np.array(faces, dtype=np.uint32)
    p = np.array(positions, dtype=np.float32)
    return MeshData(p, p.copy(), tris)


def ground_plane(half_extent: float) -> MeshData:
    """scene/shapes.rs:117-135: the XZ plane at y = 0, normal +Y."""
    e = half_extent
    p = np.array([[-e, 0, -e], [-e, 0, e], [e, 0, e], [e, 0, -e]], dtype=np.float32)
    n = np.tile(np.array([[0, 1, 0]], dtype=np.float32), (4, 1))
    uv = np.array([[0, 0], [0, 1], [1, 1], [1, 0]], dtype=np.float32)
    return MeshData(p, n, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32), uv)


def uv_sphere(segments: int, rings: int) -> MeshData:
    """A unit UV sphere: `segments` around, `rings` from pole to pole, caps
    as triangle fans — 2 * segments * (rings - 1) triangles (960 for 32x16),
    wound counter-clockwise seen from outside."""
    verts = [(0.0, 1.0, 0.0)]
    for r in range(1, rings):
        theta = np.pi * r / rings
        for s in range(segments):
            phi = 2.0 * np.pi * s / segments
            verts.append((np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)))
    verts.append((0.0, -1.0, 0.0))
    p = np.array(verts, dtype=np.float32)
    p = p / np.linalg.norm(p, axis=1, keepdims=True)
    south = len(verts) - 1
    ring = lambda r, s: 1 + (r - 1) * segments + (s % segments)  # noqa: E731
    tris = []
    for s in range(segments):
        tris.append([0, ring(1, s + 1), ring(1, s)])
    for r in range(1, rings - 1):
        for s in range(segments):
            a, b = ring(r, s), ring(r, s + 1)
            c, d = ring(r + 1, s), ring(r + 1, s + 1)
            tris.append([a, b, d])
            tris.append([a, d, c])
    for s in range(segments):
        tris.append([south, ring(rings - 1, s), ring(rings - 1, s + 1)])
    return MeshData(p, p.copy(), np.array(tris, dtype=np.uint32))


def geodesic_sphere(frequency: int) -> tuple[np.ndarray, np.ndarray]:
    """A unit geodesic sphere: every icosahedron face split into frequency^2
    triangles, vertices projected to the sphere and shared along edges —
    20 * frequency^2 triangles (32,000 at frequency 40). Returns (positions,
    indices)."""
    base, faces = _icosahedron()
    base = base.astype(np.float64)
    f = frequency
    ii, jj = [], []
    for i in range(f + 1):
        for j in range(f + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii)
    jj = np.array(jj)
    local = {(i, j): k for k, (i, j) in enumerate(zip(ii.tolist(), jj.tolist()))}
    loc_tris = []
    for i in range(f):
        for j in range(f - i):
            a, b, c = local[(i, j)], local[(i + 1, j)], local[(i, j + 1)]
            loc_tris.append([a, b, c])
            if j + 1 <= f - i - 1:
                d = local[(i + 1, j + 1)]
                loc_tris.append([b, d, c])
    loc_tris = np.array(loc_tris)
    pts, idx = [], []
    nloc = len(ii)
    for k, (a, b, c) in enumerate(faces):
        pa, pb, pc = base[a], base[b], base[c]
        # barycentric grid: i steps toward b, j toward c
        p = pa[None, :] + (pb - pa)[None, :] * (ii[:, None] / f) + (pc - pa)[None, :] * (jj[:, None] / f)
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        pts.append(p)
        idx.append(loc_tris + k * nloc)
    pts = np.concatenate(pts)
    idx = np.concatenate(idx)
    # Share vertices along face edges so the mesh is watertight.
    key = np.round(pts * 1e6).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    positions = pts[first]
    indices = inverse.reshape(-1)[idx]
    # orient counter-clockwise outward
    v0, v1, v2 = positions[indices[:, 0]], positions[indices[:, 1]], positions[indices[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(v1 - v0, v2 - v0), v0 + v1 + v2) < 0
    indices[flip] = indices[flip][:, [0, 2, 1]]
    return positions, indices.astype(np.uint32)


def vertex_normals(positions: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Area-weighted smooth normals."""
    p = positions.astype(np.float64)
    v0, v1, v2 = p[indices[:, 0]], p[indices[:, 1]], p[indices[:, 2]]
    fn = np.cross(v1 - v0, v2 - v0)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, indices[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)
