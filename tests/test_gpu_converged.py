"""Converged images: 4096 samples per pixel on the device against the CPU
oracle's 4096-sample render of the same pixels, relMSE = mean over pixels
and channels of (x - r)^2 / (r^2 + 1e-2) at most 1e-3 — on the whole
Cornell frame (at 64 x 64) and on bands of rows of the chart (1280 x 720)
and heavy (1920 x 1080) frames, where the oracle's cost allows."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SPP = 4096
REL_MSE = 1e-3


def rel_mse(x, r):
    return float(np.mean((x[..., :3].astype(np.float64) - r[..., :3]) ** 2 / (r[..., :3].astype(np.float64) ** 2 + 1e-2)))


def device_average(ctx, data, width, height, depth):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=depth)
    film = cn.Film(ctx, width, height)
    r.accumulate(scene, film, count=SPP)
    beauty = film.averages()[0]
    film.close()
    r.close()
    scene.close()
    return beauty


@pytest.mark.parametrize("name,width,height,rows", [
    ("cornell", 64, 64, (0, 64)),
    ("chart", 1280, 720, (358, 362)),
    ("heavy", 1920, 1080, (539, 541)),
])
def test_converged_render_matches_the_oracle(ctx, name, width, height, rows):
    data = workloads.build(name)
    depth = workloads.CONFIGS[name]["max_bounces"]
    dev = device_average(ctx, data, width, height, depth)[rows[0]:rows[1]]
    sums = orc.OracleScene(data).accumulate_rows(width, height, rows[0], rows[1], 0, SPP, max_bounces=depth)
    ref = sums / np.float32(SPP)
    err = rel_mse(dev, ref)
    print(f"{name} rows {rows}: relMSE {err:.3e}")
    assert err <= REL_MSE
