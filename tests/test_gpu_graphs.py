"""hipGraph replay of waves: a wave whose shape was captured before replays
the captured launch sequence with a new job table, and renders bitwise what
eager launches render — across replays with eager work in between."""

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chart(ctx):
    data = workloads.chart()
    scene = cn.Scene(ctx, data)
    yield data, scene
    scene.close()


def progressive(ctx, scene, graphs, interleave=False):
    r = cn.Renderer(ctx, capacity=96 * 54, max_bounces=12)  # one sample per wave
    r.set_graphs(graphs)
    film = cn.Film(ctx, 96, 54)
    snapshots = []
    for k in range(6):
        r.accumulate(scene, film)
        snapshots.append(film.sums()[0])
        if interleave:  # eager work between replays
            r.set_graphs(False)
            r.trace(scene, 40, 30, 100 + k)
            r.set_graphs(True)
    stats = r.graph_stats()
    film.close()
    r.close()
    return snapshots, stats


def test_replayed_waves_match_eager_waves(ctx, chart):
    _, scene = chart
    eager, _ = progressive(ctx, scene, False)
    replayed, (captures, replays) = progressive(ctx, scene, True)
    assert captures == 1 and replays == 5
    for a, b in zip(eager, replayed):
        assert np.array_equal(a, b, equal_nan=True)


def test_replays_survive_eager_work_in_between(ctx, chart):
    _, scene = chart
    eager, _ = progressive(ctx, scene, False)
    replayed, _ = progressive(ctx, scene, True, interleave=True)
    for a, b in zip(eager, replayed):
        assert np.array_equal(a, b, equal_nan=True)


def test_batches_replay_their_frames(ctx, chart):
    data, scene = chart
    views = [(64, 36, 3), (48, 48, 2), (80, 45, 4)]

    def run(graphs):
        r = cn.Renderer(ctx, capacity=64 * 36 + 48 * 48 + 80 * 45, max_bounces=12)  # one sample per view per wave
        r.set_graphs(graphs)
        films = [cn.Film(ctx, w, h) for w, h, _ in views]
        for _ in range(2):  # two progressive rounds of the same batch
            r.render_batch([(scene, f, data.camera, n) for f, (_, _, n) in zip(films, views)])
        out = [f.sums()[0] for f in films]
        stats = r.graph_stats()
        for f in films:
            f.close()
        r.close()
        return out, stats

    eager, _ = run(False)
    replayed, (captures, replays) = run(True)
    assert replays > 0 and captures < captures + replays
    for a, b in zip(eager, replayed):
        assert np.array_equal(a, b, equal_nan=True)
