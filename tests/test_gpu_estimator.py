"""The reference's closed-form estimator tests (crates/cenote/src/render/tests.rs,
wavefront.rs) run on the HIP path."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads
from cenote_amd.scene import Material, acescg_from_rec709

from reference_scenes import agreement_scene, aov_quad_scene, distant_light_scene, point_light_scene

pytestmark = pytest.mark.gpu


def _trace(ctx, data, w, h, sample=0, light_sampling=cn.LightSampling.MIS, max_bounces=8):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=max_bounces, light_sampling=light_sampling)
    img = r.trace(scene, w, h, sample)
    r.close()
    scene.close()
    return img


def _accumulate(ctx, data, w, h, spp, light_sampling=cn.LightSampling.MIS, max_bounces=8):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=max_bounces, light_sampling=light_sampling)
    film = cn.Film(ctx, w, h)
    r.accumulate(scene, film, count=spp)
    out = film.averages()
    film.close()
    r.close()
    scene.close()
    return out


def test_a_distant_light_is_analytically_exact(ctx):
    img = _trace(ctx, distant_light_scene(), 16, 16)
    err = float(np.abs(img[..., :3] - 1.0).max())
    assert err < 2e-3, err


def test_a_point_light_is_analytically_exact(ctx):
    img = _trace(ctx, point_light_scene(), 16, 16)
    err = float(np.abs(img[..., :3] - 1.0).max())
    assert err < 5e-3, err


def test_ray_offsets_hold_at_scene_scale(ctx):
    data = workloads.furnace(Material.matte((0.5, 0.5, 0.5), 0.0), center=(1e4, 0.0, 1e4), scale=1e3)
    for sample in range(4):
        img = _trace(ctx, data, 32, 32, sample, cn.LightSampling.BSDF_ONLY)
        err = float(np.abs(img[..., :3] - 0.25).max())
        assert err < 1e-3, (sample, err)


def test_the_furnace_closes_through_a_thin_lens(ctx):
    data = workloads.furnace(Material.matte((0.5, 0.5, 0.5), 0.0), aperture=0.05, focus=1.5)
    for sample in range(4):
        img = _trace(ctx, data, 32, 32, sample, cn.LightSampling.BSDF_ONLY)
        err = float(np.abs(img[..., :3] - 0.25).max())
        assert err < 1e-3, (sample, err)


def test_aovs_record_albedo_normal_and_depth(ctx):
    _, a, n, d = _accumulate(ctx, aov_quad_scene(), 64, 64, 4)
    expected = np.array(acescg_from_rec709((0.7, 0.2, 0.1)), np.float32)
    for x, y in [(32, 32), (44, 32), (32, 22)]:
        assert float(np.abs(a[y, x, :3] - expected).max()) < 1e-5
        assert float(np.abs(n[y, x, :3] - np.array([0, 0, 1], np.float32)).max()) < 1e-5
        assert abs(float(d[y, x]) - 2.0) < 1e-3
    assert float(np.abs(a[2, 2, :3] - 1.0).max()) < 1e-5
    assert float(np.abs(n[2, 2, :3]).max()) < 1e-5
    assert bool(np.isposinf(d[2, 2]))


def test_light_sampling_strategies_agree(ctx):
    means = {}
    for mode in (cn.LightSampling.MIS, cn.LightSampling.BSDF_ONLY, cn.LightSampling.NEE_ONLY):
        b, _, _, _ = _accumulate(ctx, agreement_scene(), 64, 64, 256, mode)
        means[mode] = float(b[..., :3].mean())
    mis = means[cn.LightSampling.MIS]
    assert abs(means[cn.LightSampling.BSDF_ONLY] - mis) / mis < 0.03, means
    assert abs(means[cn.LightSampling.NEE_ONLY] - mis) / mis < 0.03, means


def test_stochastic_opacity_is_weightless(ctx):
    """render/tests.rs:592-614: a half-opaque white plane in the furnace,
    BSDF-only — pass-through costs nothing, so the mean closes."""
    from cenote_amd.scene import OpenPBR

    data = workloads.furnace(OpenPBR(base_color=(1, 1, 1), specular_weight=0.0, geometry_opacity=0.5).lower())
    b, _, _, _ = _accumulate(ctx, data, 32, 32, 64, cn.LightSampling.BSDF_ONLY)
    assert abs(float(b[..., :3].mean()) - 0.5) < 0.005
