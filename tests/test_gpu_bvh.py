"""The device-built BVH is the host-built BVH: both builders make every split
with the same decision function from exact bins and split ranges by a
stable partition, so their 4-wide nodes and leaf-order triangles must be
equal bit for bit — on the benchmark scenes and on inputs chosen to stress
the corner cases (coincident centroids, degenerate and signed-zero
triangles, wide coordinate ranges, sizes around the phase boundaries)."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import _abi, workloads
from cenote_amd.scene import Material, MeshData, SceneData, identity

from bvh_nodes import contains, quantize

pytestmark = pytest.mark.gpu


def soup(positions: np.ndarray) -> SceneData:
    """A single-instance scene from (T, 3, 3) triangle corners."""
    t = positions.shape[0]
    p = np.ascontiguousarray(positions.reshape(-1, 3), dtype=np.float32)
    n = np.tile(np.array([0.0, 0.0, 1.0], np.float32), (3 * t, 1))
    idx = np.arange(3 * t, dtype=np.uint32).reshape(-1, 3)
    s = SceneData()
    s.add_instance(s.add_mesh(MeshData(p, n, idx)), identity(), Material.matte((0.5, 0.5, 0.5)))
    s.env_texels = np.ones((1, 1, 4), np.float32)
    return s


def random_soup(count, seed, scale=1.0, size=0.05):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-scale, scale, size=(count, 1, 3))
    return soup(centers + rng.normal(scale=size * scale, size=(count, 3, 3)))


CASES = {
    "cornell": workloads.cornell,
    "chart": workloads.chart,
    "one_triangle": lambda: soup(np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)),
    "coincident_5000": lambda: soup(np.tile(np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32), (5000, 1, 1))),
    "degenerate_signed_zeros": lambda: soup(np.concatenate([
        np.array([[[-0.0, 0.0, -0.0], [0.0, -0.0, 0.0], [-0.0, -0.0, 0.0]]] * 40, np.float32),
        random_soup(3000, 1).meshes[0].positions.reshape(-1, 3, 3)])),
    "soup_2048": lambda: random_soup(2048, 2),
    "soup_2049": lambda: random_soup(2049, 3),
    "soup_wide_100k": lambda: random_soup(100_000, 4, scale=1e4, size=1e-3),
    "soup_clustered_300k": lambda: random_soup(300_000, 5, scale=10.0, size=1e-4),
}


def built(ctx, data, builder, wide=False):
    scene = cn.Scene(ctx, data, bvh_builder=builder)
    nodes, tris = scene.bvh()
    info = scene.info()
    nodes8 = scene.bvh8()
    scene.close()
    return (nodes, tris, info, nodes8) if wide else (nodes, tris, info)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_and_host_builders_grow_the_same_bvh(ctx, name):
    data = CASES[name]()
    print(f"builders: {name}", flush=True)
    dn, dt, di, d8 = built(ctx, data, _abi.BVH_DEVICE, wide=True)
    hn, ht, hi, h8 = built(ctx, data, _abi.BVH_HOST, wide=True)
    assert di.bvh_builder == _abi.BVH_DEVICE and hi.bvh_builder == _abi.BVH_HOST
    assert dn.shape == hn.shape and dt.shape == ht.shape
    assert int(np.count_nonzero(dn != hn)) == 0
    assert int(np.count_nonzero(dt != ht)) == 0
    assert di.bvh_max_depth == hi.bvh_max_depth
    # the 8-wide nodes too, collapsed and quantized on the device vs the host
    assert d8.shape == h8.shape and int(np.count_nonzero(d8 != h8)) == 0
    assert di.bvh8_max_depth == hi.bvh8_max_depth
    # Every triangle appears exactly once in leaf order.
    inst_prim = dt[:, 3].astype(np.uint64) << np.uint64(32) | dt[:, 7].astype(np.uint64)
    assert np.unique(inst_prim).size == data.triangle_count


def test_heavy_scene_builds_the_same_on_the_device(ctx):
    data = workloads.heavy()
    dn, dt, di, d8 = built(ctx, data, _abi.BVH_DEVICE, wide=True)
    hn, ht, hi, h8 = built(ctx, data, _abi.BVH_HOST, wide=True)
    assert int(np.count_nonzero(dn != hn)) == 0
    assert int(np.count_nonzero(dt != ht)) == 0
    assert int(np.count_nonzero(d8 != h8)) == 0
    print(f"heavy BVH build: device {di.bvh_build_seconds:.3f} s, host {hi.bvh_build_seconds:.3f} s")


@pytest.mark.parametrize("name", ["cornell", "chart"])
def test_device_and_host_built_scenes_render_the_same_films(ctx, name):
    """The device's own quantized 4-wide nodes (never read back) walk to the
    same films as the host-quantized ones, at both widths."""
    data = workloads.build(name)
    sums = {}
    for builder in (_abi.BVH_DEVICE, _abi.BVH_HOST):
        scene = cn.Scene(ctx, data, bvh_builder=builder)
        for width in (4, 8):
            r = cn.Renderer(ctx, max_bounces=8)
            r.shape = {"bvh_width": width}
            film = cn.Film(ctx, 96, 64)
            r.accumulate(scene, film, count=2)
            sums[(builder, width)] = film.sums()
            film.close()
            r.close()
        scene.close()
    want = sums[(_abi.BVH_HOST, 4)]
    for key, got in sums.items():
        for g, w in zip(got, want):
            assert np.array_equal(g, w, equal_nan=True), key


@pytest.mark.parametrize("name", ["chart", "soup_wide_100k", "degenerate_signed_zeros"])
def test_traversal_nodes_contain_the_tree(ctx, name):
    """The 64-byte nodes the traversal reads bound the full-precision tree."""
    nodes, _, info = built(ctx, CASES[name](), _abi.BVH_HOST)
    assert info.bvh_node_bytes == 64
    assert contains(nodes, quantize(nodes))
