"""The 8-wide tree: collapsed from the same binary tree as the 4-wide one,
over the same leaves, with every quantized child box containing every
triangle below it (exact arithmetic), and walked by the traversal to the
same hits and the same films bit for bit."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads

from test_gpu_bvh import CASES

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF
LEAF = 0x80000000


def leaf_range(ref):
    return int(ref & ((1 << 27) - 1)), int((ref >> 27) & 15) + 1


def decode8(q):
    """Per node: grid origin (3), steps (3), child low / high corners
    (8, 3) in float64, child refs (8)."""
    origin = q[:, 0:3].view(np.float32).astype(np.float64)
    exps = np.stack([(q[:, 3] >> (8 * k)) & 0xFF for k in range(3)], axis=1).astype(np.int64)
    step = np.ldexp(1.0, exps - 127)

    def planes(word0, word1):
        w = np.stack([q[:, word0], q[:, word1]], axis=1)  # (N, 2)
        return np.concatenate([np.stack([(w[:, j] >> (8 * c)) & 0xFF for c in range(4)], axis=1)
                               for j in range(2)], axis=1).astype(np.float64)  # (N, 8)

    lo = np.stack([origin[:, k, None] + planes(4 + 4 * k, 5 + 4 * k) * step[:, k, None] for k in range(3)], axis=2)
    hi = np.stack([origin[:, k, None] + planes(6 + 4 * k, 7 + 4 * k) * step[:, k, None] for k in range(3)], axis=2)
    return lo, hi, q[:, 16:24]


def tri_boxes(tris):
    f = tris.view(np.float32).reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    return f.min(axis=1), f.max(axis=1)


def check_tree(nodes8, tris):
    """Every triangle in exactly one leaf reachable from the root, and every
    child's quantized box containing its whole subtree."""
    lo, hi, child = decode8(nodes8)
    tlo, thi = tri_boxes(tris)
    seen = np.zeros(len(tris), np.int64)
    leaves = []

    def subtree(index):
        """Bounds of the node's subtree; checks its children on the way."""
        box_lo, box_hi = np.full(3, np.inf), np.full(3, -np.inf)
        for c in range(8):
            ref = int(child[index, c])
            if ref == EMPTY:
                continue
            if ref & LEAF:
                first, count = leaf_range(ref)
                leaves.append((first, count))
                seen[first:first + count] += 1
                sl, sh = tlo[first:first + count].min(axis=0), thi[first:first + count].max(axis=0)
            else:
                assert ref < len(nodes8)
                sl, sh = subtree(ref)
            assert np.all(lo[index, c] <= sl) and np.all(hi[index, c] >= sh), (index, c)
            box_lo, box_hi = np.minimum(box_lo, sl), np.maximum(box_hi, sh)
        return box_lo, box_hi

    import sys

    sys.setrecursionlimit(10000)
    subtree(0)
    assert np.all(seen == 1)
    return sorted(leaves)


def bvh4_leaves(nodes):
    children = nodes[:, 24:28]
    refs = children[(children != EMPTY) & ((children & LEAF) != 0)]
    return sorted(leaf_range(int(r)) for r in refs)


@pytest.mark.parametrize("name", ["cornell", "chart", "one_triangle", "coincident_5000", "soup_2049",
                                  "soup_wide_100k"])
def test_the_8_wide_tree_covers_the_same_leaves_with_conservative_boxes(ctx, name):
    data = CASES[name]()
    scene = cn.Scene(ctx, data)
    info = scene.info()
    nodes, tris = scene.bvh()
    nodes8 = scene.bvh8()
    scene.close()
    assert info.bvh8_node_count == len(nodes8) > 0
    assert len(nodes8) < len(nodes) or len(nodes) == 1
    assert check_tree(nodes8, tris) == bvh4_leaves(nodes)


def _render(ctx, scene, width, samples, w, h, depth):
    r = cn.Renderer(ctx, max_bounces=depth)
    r.shape = {"bvh_width": width}
    film = cn.Film(ctx, w, h)
    r.accumulate(scene, film, count=samples)
    sums = film.sums()
    inst, prim, bary = r.first_hits(scene, w, h, 3)
    film.close()
    r.close()
    return sums, (inst, prim, bary)


@pytest.mark.parametrize("name,depth", [("cornell", 8), ("chart", 12), ("heavy", 16), ("heavy_field", 16)])
def test_8_wide_traversal_renders_the_same_films(ctx, name, depth):
    print(f"bvh8 films: {name}", flush=True)
    scene = cn.Scene(ctx, workloads.build(name))
    assert scene.info().bvh8_node_count > 0
    sums4, hits4 = _render(ctx, scene, 4, 3, 192, 108, depth)
    sums8, hits8 = _render(ctx, scene, 8, 3, 192, 108, depth)
    for a, b in zip(sums4, sums8):
        assert np.array_equal(a, b, equal_nan=True)
    for a, b in zip(hits4, hits8):
        assert np.array_equal(a, b)
    scene.close()
