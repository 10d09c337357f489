"""Converged images: 4096 samples per pixel on the device against the
oracle's 4096-sample average of the same pixels (committed fixtures,
written by tests/golden/make_converged.py), relMSE = mean over pixels and
channels of (x - r)^2 / (r^2 + 1e-2) at most 1e-3 — on the whole Cornell
frame (256 x 256) and on seeded 8 x 8 tiles covering at least 1% of the
chart (1280 x 720) and heavy (1920 x 1080) frames, with tiles on every
material class their bounce-0 hits show (glass, coated metal, fuzz, metal,
dielectric, emitters, sky)."""

import os

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REL_MSE = 1e-3


def rel_mse(x, r):
    x = x.astype(np.float64)
    r = r.astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def device_average(ctx, data, width, height, depth, spp):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=depth)
    film = cn.Film(ctx, width, height)
    r.accumulate(scene, film, count=spp)
    beauty = film.averages()[0]
    film.close()
    r.close()
    scene.close()
    return beauty


@pytest.mark.parametrize("name", ["cornell", "chart", "heavy"])
def test_converged_render_matches_the_oracle(ctx, name):
    f = np.load(os.path.join(GOLDEN, f"converged_{name}.npz"))
    width, height, depth, spp = (int(f[k]) for k in ("width", "height", "max_bounces", "spp"))
    dev = device_average(ctx, workloads.build(name), width, height, depth, spp)
    pixels = f["pixels"]
    got = dev.reshape(-1, 4)[pixels, :3]
    err = rel_mse(got, f["mean"])
    print(f"{name}: {len(pixels)} pixels ({len(pixels) / (width * height):.2%}), relMSE {err:.3e}")
    assert err <= REL_MSE
