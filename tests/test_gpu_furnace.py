"""The white furnace through every lobe on the HIP path, at enough samples
to resolve 1e-3: the reference's full furnace matrix (render/tests.rs:
423-488), the solid-glass furnace (:499-540), spectral interior absorption
(:541-592), coat darkening (:627-647), the specular guide ramp
(:1233-1279) and indirect colour bleeding (:1051-1116). Each furnace is
also compared with the oracle sample for sample.

The bound is 1e-3 of the sky wherever the estimator's tabulated energy
terms allow it. Five configurations keep a bias above that at any sample
count — the residual of the tabulated albedos they read (the same bias on
the oracle at 4096 spp, standard error below 1e-4): glossy r=0.05 -1.2e-3,
glossy ior=2.5 -1.2e-3, coat over diffuse -1.2e-3, the full stack -1.1e-3,
thin glass +2.6e-3. Those are held to their measured residual plus 4e-4."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads
from cenote_amd.scene import Material
from oracle import oracle as orc

from reference_scenes import FURNACE_BOUND, FURNACE_CONFIGS, color_bleed_scene, glass_sphere_scene, redness

pytestmark = pytest.mark.gpu

SKY = 0.5


def _render(ctx, data, w, h, spp, max_bounces=8, light_sampling=cn.LightSampling.MIS, first=0):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=max_bounces, light_sampling=light_sampling)
    film = cn.Film(ctx, w, h)
    r.accumulate(scene, film, count=spp)
    out = film.averages()
    film.close()
    r.close()
    scene.close()
    return out


@pytest.mark.parametrize("name", list(FURNACE_CONFIGS))
def test_openpbr_furnace_matrix_closes(ctx, name):
    print(f"furnace: {name}", flush=True)
    beauty = _render(ctx, workloads.furnace(FURNACE_CONFIGS[name]), 64, 64, 2048)[0]
    rel = (float(beauty[..., :3].mean()) - SKY) / SKY
    print(f"furnace residual {name}: {rel:+.5f}", flush=True)
    assert abs(rel) < FURNACE_BOUND[name], (name, rel)


@pytest.mark.parametrize("name", list(FURNACE_CONFIGS))
def test_furnace_samples_match_the_oracle(ctx, name):
    data = workloads.furnace(FURNACE_CONFIGS[name])
    ref = orc.OracleScene(data)
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=8)
    for sample in range(2):
        img = r.trace(scene, 32, 32, sample)
        want = ref.trace(32, 32, sample)
        close = np.isclose(img, want, rtol=1e-4, atol=1e-6).all(axis=-1).mean()
        assert close >= 0.99, (name, sample, close)
    r.close()
    scene.close()


@pytest.mark.parametrize("depth,bound", [(16, 0.015), (255, 1e-3)])
def test_the_glass_furnace_closes(ctx, depth, bound):
    """At the reference's depth cap of 16, paths cut short inside the glass
    leave about -1.1% (its bound is 1.5%); at 255 bounces the residual is
    about -5e-4."""
    beauty = _render(ctx, glass_sphere_scene(Material.glass(0.2, 1.5)), 64, 64, 1024, max_bounces=depth)[0]
    rel = (float(beauty[..., 0].mean()) - SKY) / SKY
    print(f"glass furnace residual at depth {depth}: {rel:+.5f}", flush=True)
    assert abs(rel) < bound, rel


def test_interior_absorption_is_spectral(ctx):
    from dataclasses import replace

    m = replace(Material.glass(0.2, 1.5), transmission_color=(0.4, 1.0, 1.0), transmission_depth=1.0)
    beauty = _render(ctx, glass_sphere_scene(m), 64, 64, 1024, max_bounces=255)[0]
    red, green = float(beauty[..., 0].mean()), float(beauty[..., 1].mean())
    assert abs(green - SKY) / SKY < 1e-3, green
    assert 0.2 * SKY < red < 0.9 * SKY, red


def test_coat_darkening_darkens_a_gray_base(ctx):
    from dataclasses import replace

    means = []
    for darkening in (0.0, 1.0):
        m = replace(Material.matte((0.5, 0.5, 0.5), 0.0).with_coat(1.0, 0.1), coat_darkening=darkening)
        means.append(float(_render(ctx, workloads.furnace(m), 32, 32, 256)[0][..., 0].mean()))
    ratio = means[1] / means[0]
    assert 0.45 < ratio < 0.75, (means, ratio)


def test_specular_guides_pass_through_to_what_mirrors_show(ctx):
    for roughness, expected in ((0.0, 0.035 / 0.15), (0.5, 1.0)):
        _, albedo, normal, depth = _render(ctx, workloads.furnace(Material.metal((1.0, 1.0, 1.0), roughness)),
                                           16, 16, 2)
        assert abs(float(albedo[8, 8, 0]) - 1.0) < 1e-3
        assert np.isfinite(depth[8, 8]) and depth[8, 8] > 0.0
        assert abs(float(normal[8, 8, 1]) - expected) < 1e-3, (roughness, float(normal[8, 8, 1]))


def test_indirect_light_bleeds_color(ctx):
    scene = cn.Scene(ctx, color_bleed_scene())
    r = cn.Renderer(ctx)
    film = cn.Film(ctx, 64, 64)
    r.accumulate(scene, film, count=32)
    beauty = film.sums()[0]
    assert redness(beauty, 30, 53) > 1.05 * redness(beauty, 2, 60)
    film.close()
    r.close()
    scene.close()
