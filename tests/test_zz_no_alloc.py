"""cn_accumulate_async_static, the accumulate that allocates nothing: its
film sums equal cn_accumulate_async's bit for bit, over two calls in a row
(the second on a smaller film, so it runs over the first call's leftovers
in the shared arena), and a film above the arena's maximum is refused."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums_equal(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


def _static(renderer, scene, film, count):
    from cenote_amd import _abi

    lib = _abi.library()
    _abi.check(lib.cn_accumulate_async_static(renderer.handle, scene.handle, film.handle, count))
    renderer.synchronize()


@pytest.mark.parametrize("capacity", [0, 1 << 16])
def test_static_accumulate_matches_the_allocating_one(ctx, capacity):
    import cenote_amd as cn
    from cenote_amd import workloads

    print(f"static accumulate: capacity {capacity}", flush=True)
    data = workloads.build("chart")
    scene = cn.Scene(ctx, data)
    renderer = cn.Renderer(ctx, capacity=capacity, max_bounces=12)
    # First call: the largest film the arena takes (one sample per wave);
    # second: a smaller one (several samples per wave).
    for (w, h, spp) in ((512, 512, 3), (200, 120, 9)):
        print(f"  film {w}x{h}, {spp} spp", flush=True)
        ref = cn.Film(ctx, w, h)
        got = cn.Film(ctx, w, h)
        renderer.accumulate(scene, ref, count=spp)
        _static(renderer, scene, got, spp)
        assert got.samples == spp
        want, have = ref.sums(), got.sums()
        assert _sums_equal(want, have), f"{w}x{h}: static sums differ from cn_accumulate_async"
        assert np.all(have[0][..., 3] == spp), "every pixel finishes once per sample"
        ref.close()
        got.close()
    renderer.close()
    scene.close()


def test_static_accumulate_refuses_a_film_above_its_maximum(ctx):
    import cenote_amd as cn
    from cenote_amd import _abi, workloads

    scene = cn.Scene(ctx, workloads.build("cornell"))
    renderer = cn.Renderer(ctx, max_bounces=8)
    film = cn.Film(ctx, 513, 512)
    lib = _abi.library()
    status = lib.cn_accumulate_async_static(renderer.handle, scene.handle, film.handle, 1)
    assert status == _abi.CN_E_INVALID
    assert b"exceeds the static arena" in lib.cn_last_error()
    assert film.samples == 0
    film.close()
    renderer.close()
    scene.close()
