"""The serving loop: continuous batching at sample granularity under an HBM
budget. Each request's result must be the same alone as in the mix (here:
bitwise, since a film sees its samples in index order either way) and the
same as the plain accumulate loop's average."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import _abi, serve, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(ctx):
    data = {"cornell": workloads.cornell(), "chart": workloads.chart()}
    handles = {k: cn.Scene(ctx, v) for k, v in data.items()}
    yield data, handles
    for h in handles.values():
        h.close()


MIX = [  # scene, width, height, spp, depth
    ("cornell", 64, 64, 9, 8),
    ("chart", 96, 54, 5, 12),
    ("cornell", 48, 80, 3, 5),
    ("chart", 64, 64, 7, 12),
]


def plain_average(ctx, scene, camera, w, h, spp, depth):
    r = cn.Renderer(ctx, capacity=w * h, max_bounces=depth)
    scene.set_camera(camera)
    film = cn.Film(ctx, w, h)
    for _ in range(spp):
        r.accumulate(scene, film)
    b = film.averages()[0]
    film.close()
    r.close()
    return b


def serve_all(ctx, scenes, mix, budget=1 << 30, wave_paths=0):
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=12)
    sv = serve.Server(r, budget, wave_paths)
    ids = [sv.submit(handles[name], data[name].camera, w, h, spp, max_bounces=depth)
           for name, w, h, spp, depth in mix]
    # Step, taking each result as it finishes (its film returns to the budget).
    out = [None] * len(mix)
    busy = 1
    while busy:
        busy = sv.step()
        for k, (rid, (_, w, h, spp, _)) in enumerate(zip(ids, mix)):
            st = sv.status(rid)
            if out[k] is None and st.state == _abi.REQUEST_DONE:
                assert st.samples == spp
                assert 0 <= st.admitted_seconds <= st.first_frame_seconds <= st.done_seconds
                out[k] = sv.result(rid, w, h)
                assert sv.status(rid).state == _abi.REQUEST_RELEASED
    stats = sv.stats()
    sv.close()
    r.close()
    return out, stats


def test_each_result_is_the_same_alone_and_in_the_mix(ctx, scenes):
    mixed, stats = serve_all(ctx, scenes, MIX)
    assert stats.completed == len(MIX)
    assert stats.paths == sum(w * h * spp for _, w, h, spp, _ in MIX)
    for k, req in enumerate(MIX):
        alone, _ = serve_all(ctx, scenes, [req])
        assert np.array_equal(alone[0], mixed[k], equal_nan=True), req
    data, handles = scenes
    for k, (name, w, h, spp, depth) in enumerate(MIX):
        want = plain_average(ctx, handles[name], data[name].camera, w, h, spp, depth)
        assert np.array_equal(want, mixed[k], equal_nan=True), MIX[k]
    for name in handles:
        handles[name].set_camera(data[name].camera)


def test_small_waves_and_a_tight_budget_give_the_same_results(ctx, scenes):
    mixed, _ = serve_all(ctx, scenes, MIX)
    # A budget of about one film: requests are admitted one after another.
    tight, _ = serve_all(ctx, scenes, MIX, budget=96 * 54 * 52 + 1, wave_paths=5000)
    for a, b in zip(mixed, tight):
        assert np.array_equal(a, b, equal_nan=True)


def test_drain_reports_a_budget_held_by_untaken_results(ctx, scenes):
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=8)
    sv = serve.Server(r, 64 * 64 * 52)  # one film
    a = sv.submit(handles["cornell"], data["cornell"].camera, 64, 64, 2)
    b = sv.submit(handles["cornell"], data["cornell"].camera, 64, 64, 2)
    with pytest.raises(cn.CenoteError):
        sv.drain()
    assert sv.status(a).state == _abi.REQUEST_DONE
    assert sv.status(b).state == _abi.REQUEST_QUEUED
    sv.result(a, 64, 64)
    sv.drain()
    assert sv.status(b).state == _abi.REQUEST_DONE
    sv.close()
    r.close()


def test_budget_admission_holds_requests_back(ctx, scenes):
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=8)
    film_bytes = 64 * 64 * 52
    sv = serve.Server(r, film_bytes * 2)
    ids = [sv.submit(handles["cornell"], data["cornell"].camera, 64, 64, 4) for _ in range(5)]
    seen_active = 0
    busy = 1
    while busy:
        busy = sv.step()
        st = sv.stats()
        seen_active = max(seen_active, st.active)
        assert st.hbm_used_bytes <= film_bytes * 2
        for rid in ids:
            if sv.status(rid).state == _abi.REQUEST_DONE:
                sv.result(rid, 64, 64)
    assert seen_active <= 2
    admitted = sorted(sv.status(rid).admitted_seconds for rid in ids)
    done = sorted(sv.status(rid).done_seconds for rid in ids)
    # The third request waits for the first to finish and release.
    assert admitted[2] >= done[0]
    sv.close()
    r.close()


def test_a_late_request_gets_its_first_frame_without_queueing(ctx, scenes):
    """A request arriving while a long one keeps both bulk lanes busy gets
    its first frame from the first-frame lane at once: while the long
    request still runs, and before the bulk lanes turn over more than twice
    each."""
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=8)
    sv = serve.Server(r, 1 << 30, wave_paths=64 * 64 * 4)
    long_id = sv.submit(handles["cornell"], data["cornell"].camera, 64, 64, 400)
    for _ in range(4):
        sv.step()
    late = sv.submit(handles["chart"], data["chart"].camera, 32, 32, 2, max_bounces=8)
    waves = sv.stats().waves
    steps = 0
    while sv.status(late).first_frame_seconds < 0:
        sv.step()
        steps += 1
        assert steps <= 1000
    assert sv.status(long_id).state == _abi.REQUEST_ACTIVE
    assert sv.stats().waves - waves <= 5  # the late wave and at most two per bulk lane
    sv.drain()
    sv.close()
    r.close()


def test_noise_target_stops_early_with_an_honest_estimate(ctx, scenes):
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=8)
    sv = serve.Server(r, 1 << 30, wave_paths=64 * 64 * 8)
    rid = sv.submit(handles["cornell"], data["cornell"].camera, 64, 64, 1024, rel_mse=0.05)
    sv.drain()
    st = sv.status(rid)
    assert st.state == _abi.REQUEST_DONE
    assert 2 <= st.samples < 1024
    assert 0 <= st.rel_mse_estimate <= 0.05
    got = sv.result(rid, 64, 64)
    sv.close()
    # The measured relMSE of the mean luminance against a 1024-spp render.
    ref = plain_average(ctx, handles["cornell"], data["cornell"].camera, 64, 64, 1024, 8)
    lum = lambda x: (x[..., 0] * np.float32(0.2722287) + x[..., 1] * np.float32(0.6740818)) + x[..., 2] * np.float32(0.0536895)  # noqa: E731
    measured = float(np.mean((lum(got) - lum(ref)) ** 2 / (lum(ref) ** 2 + 1e-2)))
    assert measured < 3 * 0.05
    r.close()


def test_invalid_requests_are_rejected(ctx, scenes):
    data, handles = scenes
    r = cn.Renderer(ctx, max_bounces=8)
    sv = serve.Server(r, 1 << 20)
    cam = data["cornell"].camera
    with pytest.raises(cn.CenoteError):
        sv.submit(handles["cornell"], cam, 64, 64, 0)
    with pytest.raises(cn.CenoteError):
        sv.submit(handles["cornell"], cam, 64, 64, 4, max_bounces=9)
    with pytest.raises(cn.CenoteError):
        sv.submit(handles["cornell"], cam, 4096, 4096, 4)  # film over the whole budget
    with pytest.raises(cn.CenoteError):
        sv.submit(handles["cornell"], cam, 64, 64, 4, rel_mse=-1.0)
    sv.close()
    r.close()


def test_the_serving_mix_at_full_resolution(ctx):
    """The serving workload's request mix at its stated sizes: the resident
    Cornell, chart and 2M-triangle scenes at 256^2, 1280x720 and 1920x1080,
    each scene at its own depth cap, arriving together under a budget that
    holds only a few films, in 2^22-path waves. Every result is bitwise the
    same alone as in the mix, and a noise-target request in the mix equals
    a plain render of the samples it stopped at."""
    data = {n: workloads.build(n) for n in ("cornell", "chart", "heavy")}
    handles = {n: cn.Scene(ctx, d) for n, d in data.items()}
    depth = {n: workloads.CONFIGS[n]["max_bounces"] for n in data}
    mix = [("heavy", 1920, 1080, 3), ("chart", 1280, 720, 6), ("cornell", 256, 256, 17), ("heavy", 1920, 1080, 1),
           ("cornell", 256, 256, 2), ("chart", 1280, 720, 2), ("heavy", 1280, 720, 4), ("chart", 1920, 1080, 2)]
    film_1080p = 1920 * 1080 * 56
    budget = 3 * film_1080p

    def run(requests, noisy=None):
        r = cn.Renderer(ctx, max_bounces=16)
        sv = serve.Server(r, budget, 1 << 22)
        ids = [sv.submit(handles[n], data[n].camera, w, h, spp, max_bounces=depth[n]) for n, w, h, spp in requests]
        nid = None
        if noisy is not None:
            n, w, h = noisy
            nid = sv.submit(handles[n], data[n].camera, w, h, 256, rel_mse=1e-2, max_bounces=depth[n])
        out, noise = [None] * len(ids), None
        busy = 1
        while busy:
            busy = sv.step()
            assert sv.stats().hbm_used_bytes <= budget
            for k, (rid, (_, w, h, spp)) in enumerate(zip(ids, requests)):
                if out[k] is None and sv.status(rid).state == _abi.REQUEST_DONE:
                    assert sv.status(rid).samples == spp
                    out[k] = sv.result(rid, w, h)
            if nid is not None and noise is None and sv.status(nid).state == _abi.REQUEST_DONE:
                noise = (sv.status(nid).samples, sv.result(nid, noisy[1], noisy[2]))
        sv.close()
        r.close()
        return out, noise

    print("serving mix: all requests together", flush=True)
    mixed, (noise_samples, noise_result) = run(mix, noisy=("cornell", 256, 256))
    for k, req in enumerate(mix):
        print(f"serving mix: {req} alone", flush=True)
        alone, _ = run([req])
        assert np.array_equal(alone[0], mixed[k], equal_nan=True), req
    assert 2 <= noise_samples <= 256
    want = plain_average(ctx, handles["cornell"], data["cornell"].camera, 256, 256, noise_samples, depth["cornell"])
    assert np.array_equal(want, noise_result, equal_nan=True)
    for h in handles.values():
        h.close()
