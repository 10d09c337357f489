"""Fractional opacity and the camera mask on the HIP path
(intersect.slang:64-88, trace_shadow.slang:41-71, scene/mod.rs:54-57),
against the same closed forms as the oracle and against the oracle itself
sample for sample: the per-crossing trials and the fixed-point shadow
transmittance are order-free, so the BVH4 walk and the oracle's binary-tree
walk must make the very same decisions."""

import math

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads
from cenote_amd.scene import Material
from oracle import oracle as orc

from reference_scenes import invisible_emitter_scene, stacked_planes_scene, translucent_scene

pytestmark = pytest.mark.gpu


def _trace(ctx, data, w, h, sample=0, light_sampling=cn.LightSampling.MIS, max_bounces=8):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=max_bounces, light_sampling=light_sampling)
    img = r.trace(scene, w, h, sample)
    r.close()
    scene.close()
    return img


@pytest.mark.parametrize("planes,expected", [(1, 0.7), (20, 0.7 ** 20), (25, 0.7 ** 25), (26, 0.0), (30, 0.0)])
def test_stacked_planes_attenuate_every_crossing(ctx, planes, expected):
    print(f"stacked planes: {planes}", flush=True)
    data = stacked_planes_scene(planes, 0.3)
    ref = orc.OracleScene(data)
    for sample in range(2):
        img = _trace(ctx, data, 16, 16, sample)
        if expected == 0.0:
            assert np.all(img[..., :3] == 0.0)
        else:
            assert np.allclose(img[..., :3], expected, rtol=2e-6, atol=0.0), (img[..., 0].min(), expected)
        assert np.array_equal(img, ref.trace(16, 16, sample))


def test_stochastic_opacity_is_per_sample_exact(ctx):
    """render/tests.rs:593-614, per sample and per pixel."""
    data = workloads.furnace(Material.matte((1.0, 1.0, 1.0), 0.0).with_opacity(0.5))
    for sample in range(4):
        img = _trace(ctx, data, 32, 32, sample, cn.LightSampling.BSDF_ONLY)
        assert np.all(np.abs(img[..., :3] - 0.5) < 1e-3), sample


def test_opacity_first_hits_match_the_oracle(ctx):
    """Four half-opaque planes seen straight on: the same instance on every
    pixel as the oracle, and the nearest-committing-crossing shares."""
    data = stacked_planes_scene(4, 0.5, camera_visible=True)
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=8)
    ref = orc.OracleScene(data)
    for sample in range(2):
        inst, prim, _ = r.first_hits(scene, 128, 128, sample)
        ref_inst, ref_prim, _ = ref.first_hits(128, 128, sample)
        assert np.array_equal(inst, ref_inst) and np.array_equal(prim, ref_prim), sample
        for j in range(4):
            p = 0.5 ** (j + 1)
            assert abs(np.mean(inst == 4 - j) - p) < 4 * math.sqrt(p * (1 - p) / inst.size)
    r.close()
    scene.close()


def test_an_invisible_emitter_lights_without_appearing(ctx):
    """scene/prep.rs:467-558, and the same pixels as the oracle."""
    size = 64
    images = {}
    for visible in (True, False):
        data = invisible_emitter_scene(visible)
        img = _trace(ctx, data, size, size)
        ref = orc.OracleScene(data).trace(size, size, 0)
        assert np.isclose(img, ref, rtol=1e-4, atol=1e-6).all(axis=-1).mean() >= 0.999
        images[visible] = img
    assert images[True][size // 2, size // 2, 0] > 5.0
    assert images[False][size // 2, size // 2, 0] < 0.5
    for img in images.values():
        assert img[size - 2, size // 2, 0] > 0.01


def test_translucent_scene_matches_the_oracle_per_sample(ctx):
    """Half-opaque sphere, a translucent pane across the light and a
    camera-invisible emitter in a lit scene, every light-sampling mode:
    per-sample matched parity with the oracle (transcendental ULPs aside)."""
    data = translucent_scene()
    ref = orc.OracleScene(data)
    for mode in (cn.LightSampling.MIS, cn.LightSampling.BSDF_ONLY, cn.LightSampling.NEE_ONLY):
        for sample in range(3):
            print(f"translucent: mode {int(mode)} sample {sample}", flush=True)
            img = _trace(ctx, data, 96, 64, sample, mode)
            want = ref.trace(96, 64, sample, light_sampling=int(mode))
            close = np.isclose(img, want, rtol=1e-4, atol=1e-6).all(axis=-1).mean()
            assert close >= 0.99, (int(mode), sample, close)
