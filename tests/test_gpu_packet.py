"""Camera rays walk the tree as wave packets, except in counting waves, which
walk them lane by lane. The closest hit does not depend on which nodes a walk
visits, so both walks render bitwise the same film — including waves whose
rays' direction signs differ on some axis (the axis-aligned Cornell camera's
centre row and column, the heavy field view's diagonal), the thin-lens chart,
translucent and camera-invisible surfaces, and stacked fractional-opacity
planes seen head on."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads

from reference_scenes import invisible_emitter_scene, stacked_planes_scene, translucent_scene

SCENES = {
    "translucent": translucent_scene,
    "invisible_emitter": lambda: invisible_emitter_scene(False),
    "stacked_planes": lambda: stacked_planes_scene(12, 0.5, camera_visible=True),
}

pytestmark = pytest.mark.gpu


def _film(ctx, scene, width, height, samples, depth, counting):
    r = cn.Renderer(ctx, max_bounces=depth)
    r.set_counters(counting)
    film = cn.Film(ctx, width, height)
    r.accumulate(scene, film, count=samples)
    sums = film.sums()
    film.close()
    r.close()
    return sums


@pytest.mark.parametrize("name,width,height,samples,depth", [
    ("cornell", 96, 96, 4, 8),
    ("chart", 160, 90, 3, 12),
    ("heavy", 192, 108, 2, 16),
    ("heavy_field", 192, 108, 2, 16),
    ("translucent", 64, 64, 4, 8),
    ("invisible_emitter", 64, 64, 4, 8),
    ("stacked_planes", 48, 48, 4, 8),
])
def test_packet_walk_renders_the_per_lane_film(ctx, name, width, height, samples, depth):
    print(f"packet: {name} {width}x{height} spp {samples}", flush=True)
    data = SCENES[name]() if name in SCENES else workloads.build(name)
    scene = cn.Scene(ctx, data)
    try:
        packets = _film(ctx, scene, width, height, samples, depth, counting=False)
        lanes = _film(ctx, scene, width, height, samples, depth, counting=True)
    finally:
        scene.close()
    for k, (a, b) in enumerate(zip(packets, lanes)):
        assert np.array_equal(a, b, equal_nan=True), (name, k, int(np.sum(a != b)))
