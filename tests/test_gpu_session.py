"""The progressive session (render/session.rs): frames arrive while the
thread accumulates, each the exact average of its sample count; camera
changes and scene edits restart accumulation; a rejected edit is posted
and the last good scene keeps rendering; a thread failure surfaces."""

import dataclasses
import time

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import scenefile
from cenote_amd.session import Session

pytestmark = pytest.mark.gpu

SCENE = '''(version: 1, ops: [
    Camera((name: "main", position: Some((0.0, 1.0, 4.0)), look_at: Some((0.0, 0.5, 0.0)))),
    Mesh((name: "floor", source: Some(Inline(
        positions: [(-2.0, 0.0, -2.0), (-2.0, 0.0, 2.0), (2.0, 0.0, 2.0), (2.0, 0.0, -2.0)],
        triangles: [(0, 1, 2), (0, 2, 3)])))),
    Mesh((name: "lamp", source: Some(Inline(
        positions: [(-0.5, 2.0, -0.5), (0.5, 2.0, -0.5), (0.5, 2.0, 0.5), (-0.5, 2.0, 0.5)],
        triangles: [(0, 1, 2), (0, 2, 3)])))),
    Material((name: "clay", base_color: Some(Constant((0.7, 0.4, 0.3))))),
    Material((name: "light", emission_luminance: Some(8.0))),
    Instance((name: "floor", mesh: Some("floor"), material: Some("clay"))),
    Instance((name: "lamp", mesh: Some("lamp"), material: Some("light"))),
])'''

W, H = 64, 48


def description():
    d = scenefile.Description()
    d.apply(scenefile.from_ron(SCENE))
    return d


def wait_frame(session, at_least, timeout=30.0):
    deadline = time.time() + timeout
    while time.time() < deadline:
        session.check()
        f = session.take_frame()
        if f is not None and f.samples >= at_least:
            return f
        time.sleep(0.005)
    raise AssertionError(f"no frame with {at_least} samples in {timeout} s")


def plain_average(ctx, d, camera, samples):
    data, _ = scenefile.lower(d)
    data.camera = camera
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=8)
    film = cn.Film(ctx, W, H)
    r.accumulate(scene, film, count=samples)
    out = r.resolve(film)[0]
    film.close()
    r.close()
    scene.close()
    return out


def test_frames_are_exact_averages_of_their_samples(ctx):
    d = description()
    s = Session(ctx, d, W, H)
    f = wait_frame(s, 4)
    s.close()
    assert (f.width, f.height) == (W, H) and f.sample_time > 0
    data, _ = scenefile.lower(d)
    assert np.array_equal(f.beauty, plain_average(ctx, d, data.camera, f.samples))


def test_a_camera_change_restarts_accumulation(ctx):
    d = description()
    s = Session(ctx, d, W, H)
    wait_frame(s, 8)
    data, _ = scenefile.lower(d)
    moved = dataclasses.replace(data.camera, position=(1.0, 1.5, 3.5))
    s.set_camera(moved)
    time.sleep(0.05)
    s.take_frame()
    f = wait_frame(s, 2)
    s.close()
    assert np.array_equal(f.beauty, plain_average(ctx, d, moved, f.samples))


def test_edits_land_and_bad_edits_are_posted(ctx):
    d = description()
    s = Session(ctx, d, W, H)
    wait_frame(s, 2)
    s.apply(scenefile.from_ron('(version: 1, ops: [Instance((name: "x", mesh: Some("nope"), material: Some("clay")))])'))
    deadline = time.time() + 10
    err = None
    while err is None and time.time() < deadline:
        err = s.take_edit_error()
        time.sleep(0.01)
    assert isinstance(err, scenefile.SceneFileError)
    wait_frame(s, 2)  # still rendering the last good scene
    edit = scenefile.from_ron('(version: 1, ops: [Material((name: "clay", base_color: Some(Constant((0.1, 0.9, 0.1)))))])')
    s.apply(edit)
    time.sleep(0.1)
    s.take_frame()
    f = wait_frame(s, 2)
    s.close()
    d.apply(edit)
    data, _ = scenefile.lower(d)
    assert np.array_equal(f.beauty, plain_average(ctx, d, data.camera, f.samples))


def test_a_render_thread_failure_surfaces(ctx):
    d = description()
    s = Session(ctx, d, W, H)
    wait_frame(s, 1)
    data, _ = scenefile.lower(d)
    s.set_camera(dataclasses.replace(data.camera, position=data.camera.look_at))  # degenerate: rejected by the library
    deadline = time.time() + 10
    with pytest.raises(cn.CenoteError):
        while time.time() < deadline:
            s.check()
            time.sleep(0.01)
    s.close()  # the failure was surfaced once; closing is clean
