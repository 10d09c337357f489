"""Camera sampling and light-sampling strategies on the HIP path: the
reference's jitter, aperture, strategy-agreement and ragged-size tests
(wavefront.rs:957-1031, :1148-1226, :1267-1356, render/tests.rs:164-184),
each image also matched sample for sample with the oracle, and the strategy
agreement checked region by region — 16 tiles of 8x8 pixels, each within
five standard errors (from 8 independent sample batches) plus 1% — so a
MIS weight that is wrong over one part of the frame cannot hide in the
frame mean."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads
from oracle import oracle as orc

from reference_scenes import env_strategy_scene, mixed_pixels, silhouette_scene, strategy_scene

pytestmark = pytest.mark.gpu


def _images(ctx, data, samples, w=32, h=32, capacity=4096):
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, capacity=capacity, max_bounces=8)
    out = [r.trace(scene, w, h, s) for s in range(samples)]
    r.close()
    scene.close()
    return out


def test_camera_jitter_mixes_edge_pixels(ctx):
    data = silhouette_scene()
    images = _images(ctx, data, 16)
    mixed, sky = mixed_pixels(images)
    assert not sky[:, 31, 16].any(), "the plane-interior pixel saw the sky"
    assert mixed > 0
    ref = orc.OracleScene(data)
    assert mixed == mixed_pixels([ref.trace(32, 32, s) for s in range(16)])[0]


def test_an_open_aperture_blurs_out_of_focus_silhouettes(ctx):
    pinhole = mixed_pixels(_images(ctx, silhouette_scene(), 16))[0]
    data = silhouette_scene(aperture=0.4, focus=2.0)
    lensed = _images(ctx, data, 16)
    blurred = mixed_pixels(lensed)[0]
    assert pinhole > 0
    assert blurred > 2 * pinhole, (blurred, pinhole)
    ref = orc.OracleScene(data)
    for s in (0, 7):
        close = np.isclose(lensed[s], ref.trace(32, 32, s), rtol=1e-4, atol=1e-6).all(axis=-1).mean()
        assert close >= 0.99, (s, close)


def _tile_means(image):
    """Per-tile means of an (H, W, 4) image over 8x8-pixel tiles, RGB summed."""
    rgb = image[..., :3].astype(np.float64).sum(axis=-1)
    h, w = rgb.shape
    return rgb.reshape(h // 8, 8, w // 8, 8).mean(axis=(1, 3)).ravel()


@pytest.mark.parametrize("scene_fn", [strategy_scene, env_strategy_scene], ids=["emitter", "environment"])
def test_light_sampling_strategies_agree_region_by_region(ctx, scene_fn):
    scene = cn.Scene(ctx, scene_fn())
    batches, per = 8, 512
    stats = {}
    for mode in (cn.LightSampling.MIS, cn.LightSampling.BSDF_ONLY, cn.LightSampling.NEE_ONLY):
        r = cn.Renderer(ctx, max_bounces=8, light_sampling=mode)
        film = cn.Film(ctx, 32, 32)
        prev = np.zeros((32, 32, 4))
        tiles = []
        for _ in range(batches):  # consecutive sample ranges of one film: independent batches
            r.accumulate(scene, film, count=per)
            total = film.sums()[0].astype(np.float64)
            tiles.append(_tile_means((total - prev) / per))
            prev = total
        tiles = np.array(tiles)
        stats[int(mode)] = (tiles.mean(axis=0), tiles.std(axis=0, ddof=1) / np.sqrt(batches))
        film.close()
        r.close()
    scene.close()
    mis, mis_se = stats[int(cn.LightSampling.MIS)]
    assert mis.mean() > 0.01
    for mode in (cn.LightSampling.BSDF_ONLY, cn.LightSampling.NEE_ONLY):
        value, se = stats[int(mode)]
        # the reference's whole-frame bound
        assert abs(value.mean() - mis.mean()) / mis.mean() < 0.03, (int(mode), value.mean(), mis.mean())
        # and every region
        bound = 5.0 * np.sqrt(se ** 2 + mis_se ** 2) + 0.01 * np.abs(mis) + 1e-6
        worst = np.abs(value - mis) / bound
        assert np.all(worst < 1.0), (int(mode), float(worst.max()), value, mis)


@pytest.mark.parametrize("size", [(33, 17), (1, 1), (7, 3), (300, 1)])
def test_ragged_dimensions_cover_every_pixel(ctx, size):
    w, h = size
    data = workloads.chart()
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=12)
    img = r.trace(scene, w, h, 0)
    assert np.all(img[..., 3] == 1.0)
    assert np.all(np.isfinite(img)) and np.all(img[..., :3] >= 0.0)
    ref = orc.OracleScene(data).trace(w, h, 0, max_bounces=12)
    assert np.isclose(img, ref, rtol=1e-4, atol=1e-6).all(axis=-1).mean() >= 0.95
    r.close()
    scene.close()
