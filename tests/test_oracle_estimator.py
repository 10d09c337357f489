"""Pins for the oracle's estimator (the per-path restatement of
shaders/raygen.slang .. trace_shadow.slang): the reference's closed-form
tests of crates/cenote/src/render/tests.rs and wavefront.rs, at sizes the
CPU finishes in seconds."""

import numpy as np
import pytest

from cenote_amd import workloads
from cenote_amd.scene import Material, OpenPBR, acescg_from_rec709
from oracle import oracle as orc

from reference_scenes import (FURNACE_BOUND, FURNACE_CONFIGS, agreement_scene, aov_quad_scene, color_bleed_scene,
                              distant_light_scene, glass_sphere_scene, point_light_scene, redness)

BSDF_ONLY, NEE_ONLY, MIS = 1, 2, 0


def test_lambert_furnace_closes_per_sample():
    """render/tests.rs:185-203: BSDF-only, every sample = albedo x sky."""
    scene = orc.OracleScene(workloads.furnace(Material.matte((1.0, 1.0, 1.0), 0.0)))
    for sample in range(4):
        img = scene.trace(32, 32, sample, light_sampling=BSDF_ONLY)
        assert np.allclose(img[..., :3], 0.5, atol=1e-3)
        assert np.all(img[..., 3] == 1.0)


def test_rough_diffuse_furnace_mean():
    """render/tests.rs:205-213: EON at roughness 1 closes to 0.005."""
    scene = orc.OracleScene(workloads.furnace(Material.matte((1.0, 1.0, 1.0), 1.0)))
    b, _, _, _ = scene.accumulate(32, 32, 0, 64)
    assert abs(b[..., 0].mean() / 64 - 0.5) < 0.005


def test_ray_offsets_hold_at_scene_scale():
    """render/tests.rs:226-253: the plane 1e4 m out and 1000x larger."""
    data = workloads.furnace(Material.matte((0.5, 0.5, 0.5), 0.0), center=(1e4, 0.0, 1e4), scale=1e3)
    scene = orc.OracleScene(data)
    for sample in range(4):
        img = scene.trace(32, 32, sample, light_sampling=BSDF_ONLY)
        assert np.allclose(img[..., :3], 0.25, atol=1e-3)


def test_the_furnace_closes_through_a_thin_lens():
    """render/tests.rs:265-286"""
    data = workloads.furnace(Material.matte((0.5, 0.5, 0.5), 0.0), aperture=0.05, focus=1.5)
    scene = orc.OracleScene(data)
    for sample in range(4):
        img = scene.trace(32, 32, sample, light_sampling=BSDF_ONLY)
        assert np.allclose(img[..., :3], 0.25, atol=1e-3)


def test_a_distant_light_is_analytically_exact():
    """render/tests.rs:352-374"""
    img = orc.OracleScene(distant_light_scene()).trace(16, 16, 0)
    assert np.allclose(img[..., :3], 1.0, atol=2e-3)


def test_a_point_light_is_analytically_exact():
    """render/tests.rs:383-408"""
    img = orc.OracleScene(point_light_scene()).trace(16, 16, 0)
    assert np.allclose(img[..., :3], 1.0, atol=5e-3)


@pytest.mark.parametrize("name", list(FURNACE_CONFIGS))
def test_openpbr_furnace_matrix(name):
    """render/tests.rs:422-488: a white material of any construction returns
    the sky: within 1e-3 wherever the tabulated energy terms allow, within
    the measured table residual elsewhere (FURNACE_BOUND)."""
    scene = orc.OracleScene(workloads.furnace(FURNACE_CONFIGS[name]))
    b, _, _, _ = scene.accumulate(32, 32, 0, 1024)
    rel = (b[..., :3].mean() / 1024 - 0.5) / 0.5
    assert abs(rel) < FURNACE_BOUND[name], (name, rel)


@pytest.mark.parametrize("depth,bound", [(16, 0.015), (255, 1e-3)])
def test_the_glass_furnace_closes(depth, bound):
    """render/tests.rs:499-540: a closed rough-glass sphere. At the
    reference's depth cap of 16 the paths cut short inside the glass leave
    -1.1% (the reference's bound is 1.5%); at 255 bounces the residual is
    -5e-4, inside 1e-3."""
    scene = orc.OracleScene(glass_sphere_scene(Material.glass(0.2, 1.5)))
    b, _, _, _ = scene.accumulate(32, 32, 0, 512, max_bounces=depth)
    rel = (b[..., 0].mean() / 512 - 0.5) / 0.5
    assert abs(rel) < bound, rel


def test_interior_absorption_is_spectral():
    """render/tests.rs:541-592: red absorbed along interior chords, green
    (absorption-free) still closes."""
    from dataclasses import replace

    m = replace(Material.glass(0.2, 1.5), transmission_color=(0.4, 1.0, 1.0), transmission_depth=1.0)
    b, _, _, _ = orc.OracleScene(glass_sphere_scene(m)).accumulate(32, 32, 0, 256, max_bounces=255)
    red, green = b[..., 0].mean() / 256, b[..., 1].mean() / 256
    assert abs(green - 0.5) / 0.5 < 2e-3, green
    assert 0.2 * 0.5 < red < 0.9 * 0.5, red


def test_coat_darkening_darkens_a_gray_base():
    """render/tests.rs:627-647: darkening 0 -> 1 under a 1.6-IOR coat over a
    0.5 Lambert base lands near the spec's 0.52."""
    from dataclasses import replace

    means = []
    for darkening in (0.0, 1.0):
        m = replace(Material.matte((0.5, 0.5, 0.5), 0.0).with_coat(1.0, 0.1), coat_darkening=darkening)
        b, _, _, _ = orc.OracleScene(workloads.furnace(m)).accumulate(16, 16, 0, 32)
        means.append(b[..., 0].mean() / 32)
    assert 0.45 < means[1] / means[0] < 0.75, means


def test_specular_guides_pass_through_to_what_mirrors_show():
    """render/tests.rs:1233-1279: a mirror records 0.035/0.15 of its normal
    (the roughness floor into the guide ramp), roughness 0.5 all of it."""
    for roughness, expected in ((0.0, 0.035 / 0.15), (0.5, 1.0)):
        scene = orc.OracleScene(workloads.furnace(Material.metal((1.0, 1.0, 1.0), roughness)))
        _, a, n, d = scene.accumulate(16, 16, 0, 2)
        assert abs(a[8, 8, 0] / 2 - 1.0) < 1e-3
        assert np.isfinite(d[8, 8]) and d[8, 8] > 0.0
        assert abs(n[8, 8, 1] / 2 - expected) < 1e-3, (roughness, n[8, 8, 1] / 2)


def test_indirect_light_bleeds_color():
    """render/tests.rs:1051-1116: floor beside the terracotta sphere is
    redder than the far corner."""
    b, _, _, _ = orc.OracleScene(color_bleed_scene()).accumulate(64, 64, 0, 32)
    assert redness(b, 30, 53) > 1.05 * redness(b, 2, 60)


def test_aovs_record_albedo_normal_and_depth():
    """render/tests.rs:1117-1221"""
    scene = orc.OracleScene(aov_quad_scene())
    b, a, n, d = scene.accumulate(64, 64, 0, 4)
    a, n, d = a / 4, n / 4, d / 4
    expected = np.array(acescg_from_rec709((0.7, 0.2, 0.1)), np.float32)
    for x, y in [(32, 32), (44, 32), (32, 22)]:
        assert np.allclose(a[y, x, :3], expected, atol=1e-5)
        assert np.allclose(n[y, x, :3], [0.0, 0.0, 1.0], atol=1e-5)
        assert abs(d[y, x] - 2.0) < 1e-3
    assert np.allclose(a[2, 2, :3], 1.0, atol=1e-5)
    assert np.allclose(n[2, 2, :3], 0.0, atol=1e-5)
    assert np.isposinf(d[2, 2])


def test_light_sampling_strategies_agree():
    """wavefront.rs:1246-1557: MIS, BSDF-only and NEE-only converge to the
    same image (here: the same mean over the frame, within 3%)."""
    scene = orc.OracleScene(agreement_scene())
    means = {}
    for mode in (MIS, BSDF_ONLY, NEE_ONLY):
        b, _, _, _ = scene.accumulate(32, 32, 0, 256, light_sampling=mode)
        means[mode] = b[..., :3].mean() / 256
    assert abs(means[BSDF_ONLY] - means[MIS]) / means[MIS] < 0.03, means
    assert abs(means[NEE_ONLY] - means[MIS]) / means[MIS] < 0.03, means


def test_threads_do_not_change_the_image():
    scene = orc.OracleScene(workloads.cornell())
    a = scene.trace(64, 64, 3, threads=1)
    b = scene.trace(64, 64, 3, threads=7)
    assert np.array_equal(a, b)


def test_every_pixel_finishes_exactly_once():
    scene = orc.OracleScene(workloads.cornell())
    img = scene.trace(64, 64, 0)
    assert np.all(img[..., 3] == 1.0)


def test_camera_basis_is_the_glam_basis():
    """scene/mod.rs:448-461: forward unit, right = fwd x up scaled by
    tan(vfov/2) * aspect, up = right x fwd scaled by tan(vfov/2)."""
    import ctypes as C

    from cenote_amd._abi import Camera, f3

    cam = Camera(f3(0, 5.5, 11), f3(0, 0, 0), f3(0, 1, 0), 40.0, 0.0, 0.0)
    out = (C.c_float * 9)()
    orc.lib().orc_camera_basis(C.cast(C.pointer(cam), C.c_void_p), 16 / 9, out)
    right, up, fwd = np.array(out[0:3]), np.array(out[3:6]), np.array(out[6:9])
    f = np.array([0, -5.5, -11.0])
    f /= np.linalg.norm(f)
    r = np.cross(f, [0, 1, 0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    t = np.tan(np.radians(20.0))
    assert np.allclose(fwd, f, atol=1e-6)
    assert np.allclose(right, r * t * 16 / 9, atol=1e-6)
    assert np.allclose(up, u * t, atol=1e-6)


def test_band_accumulation_matches_full_frames():
    """The pixel-parallel band accumulation folds the same samples in the
    same order as whole-frame accumulation."""
    from cenote_amd import workloads as wl

    s = orc.OracleScene(wl.cornell())
    full = s.accumulate(24, 20, 0, 3)[0]
    band = s.accumulate_rows(24, 20, 5, 12, 0, 3)
    assert np.array_equal(band, full[5:12])
