"""Pins for the oracle's camera sampling and light-sampling strategies, the
reference's behaviour tests restated: jitter that mixes silhouette pixels
and never leaves a pixel's footprint (wavefront.rs:957-1031), an open
aperture that blurs out-of-focus silhouettes (wavefront.rs:1148-1226),
MIS / BSDF-only / NEE-only agreement on an emitter scene and on the
environment (wavefront.rs:1267-1356, :1510-1557), and ragged target sizes
(render/tests.rs:164-184)."""

import numpy as np
import pytest

from cenote_amd import workloads
from oracle import oracle as orc

from reference_scenes import env_strategy_scene, mixed_pixels, silhouette_scene, strategy_scene


def test_camera_jitter_mixes_edge_pixels():
    scene = orc.OracleScene(silhouette_scene())
    images = [scene.trace(32, 32, s) for s in range(16)]
    mixed, sky = mixed_pixels(images)
    assert not sky[:, 31, 16].any(), "the plane-interior pixel saw the sky"
    assert mixed > 0


def test_an_open_aperture_blurs_out_of_focus_silhouettes():
    pinhole = mixed_pixels([orc.OracleScene(silhouette_scene()).trace(32, 32, s) for s in range(16)])[0]
    lens = orc.OracleScene(silhouette_scene(aperture=0.4, focus=2.0))
    blurred = mixed_pixels([lens.trace(32, 32, s) for s in range(16)])[0]
    assert pinhole > 0
    assert blurred > 2 * pinhole, (blurred, pinhole)


@pytest.mark.parametrize("scene_fn", [strategy_scene, env_strategy_scene], ids=["emitter", "environment"])
def test_light_sampling_strategies_agree(scene_fn):
    scene = orc.OracleScene(scene_fn())
    means = {}
    for mode in (0, 1, 2):
        b, _, _, _ = scene.accumulate(32, 32, 0, 256, light_sampling=mode)
        means[mode] = float(b[..., :3].astype(np.float64).sum() / (256 * 32 * 32))
    assert means[0] > 0.01
    for mode in (1, 2):
        assert abs(means[mode] - means[0]) / means[0] < 0.03, means


@pytest.mark.parametrize("size", [(33, 17), (1, 1), (7, 3), (300, 1)])
def test_ragged_dimensions_cover_every_pixel(size):
    w, h = size
    img = orc.OracleScene(workloads.chart()).trace(w, h, 0, max_bounces=12)
    assert np.all(img[..., 3] == 1.0)
    assert np.all(np.isfinite(img)) and np.all(img[..., :3] >= 0.0)
