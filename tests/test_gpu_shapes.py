"""Launch shapes change speed, never results: every shade_surface variant and
grid size renders bitwise the same film (the stage kernels are per-path
pure functions of the path's own state, and queue order never reaches a
slot's add order)."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import _abi, autotune, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chart(ctx):
    data = workloads.chart()
    scene = cn.Scene(ctx, data)
    yield data, scene
    scene.close()


def render_sums(ctx, scene, shape, width=128, height=72, samples=3):
    r = cn.Renderer(ctx, max_bounces=12)
    if shape:
        r.shape = shape
    film = cn.Film(ctx, width, height)
    r.accumulate(scene, film, count=samples)
    sums = film.sums()
    film.close()
    r.close()
    return sums


@pytest.mark.parametrize("variant", [_abi.SHADE_1024_LDS, _abi.SHADE_512_LDS, _abi.SHADE_256_LDS,
                                     _abi.SHADE_256_GLOBAL, _abi.SHADE_512_LDS_SORTED,
                                     _abi.SHADE_256_LDS_SORTED, _abi.SHADE_256_GLOBAL_SORTED])
@pytest.mark.parametrize("trace_blocks,shade_blocks,light_blocks,width", [(7, 3, 5, 4), (2048, 512, 1024, 4),
                                                                           (7, 3, 5, 8), (2048, 512, 1024, 8)])
def test_every_launch_shape_renders_the_same_film(ctx, chart, variant, trace_blocks, shade_blocks, light_blocks,
                                                  width):
    _, scene = chart
    want = render_sums(ctx, scene, None)
    got = render_sums(ctx, scene, {"trace_blocks": trace_blocks, "shade_variant": variant,
                                   "shade_blocks": shade_blocks, "light_blocks": light_blocks, "bvh_width": width})
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)


def test_invalid_shapes_are_rejected(ctx):
    r = cn.Renderer(ctx, max_bounces=4)
    before = r.shape
    for bad in ({"trace_blocks": 0}, {"shade_variant": 9}, {"shade_blocks": 70000}, {"light_blocks": 0},
                {"bvh_width": 2}, {"bvh_width": 16}):
        with pytest.raises(cn.CenoteError):
            r.shape = bad
    assert r.shape == before
    r.close()


def test_autotuner_sets_the_fastest_candidate_and_keeps_results(ctx, chart):
    _, scene = chart
    _, cus = ctx.describe()
    r = cn.Renderer(ctx, max_bounces=12)
    candidates = list(autotune.candidate_shapes(cus, trace_per_cu=(4, 8)))
    result = autotune.tune(r, scene, 128, 72, samples=2, repeats=1, candidates=candidates)
    assert r.shape == result["shape"]
    assert result["shape"] in candidates
    assert len(result["timings"]) == len(candidates)
    assert result["seconds"] == min(t["seconds"] for t in result["timings"])
    film = cn.Film(ctx, 128, 72)
    r.accumulate(scene, film, count=3)
    got = film.sums()
    film.close()
    r.close()
    want = render_sums(ctx, scene, None)
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)


def test_capacity_tuner_picks_a_measured_capacity(ctx, chart):
    _, scene = chart
    result = autotune.tune_capacity(ctx, scene, 128, 72, samples=4, max_bounces=12,
                                    capacities=(128 * 72, 2 * 128 * 72, 4 * 128 * 72), repeats=1)
    assert result["capacity"] in (128 * 72, 2 * 128 * 72, 4 * 128 * 72)
    assert result["seconds"] == min(t["seconds"] for t in result["timings"])
