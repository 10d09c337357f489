"""The headless render command (crates/cenote-cli/src/main.rs:30-71):
argument ranges and scene defaults without a GPU; a full render to a
multi-layer EXR on the GPU."""

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from cenote_amd import cli

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("argv", [["render", "--spp", "0"], ["render", "--depth", "256"], ["render", "--width", "0"]])
def test_flags_outside_their_ranges_are_rejected(argv):
    with pytest.raises(SystemExit):
        cli.build_parser().parse_args(argv)


def test_an_unknown_scene_is_an_error_before_any_device_work(capsys):
    assert cli.main(["render", "no_such_scene"]) == 2
    assert "unknown scene" in capsys.readouterr().err
    assert cli.main(["render", "no_such_scene.ron"]) == 2
    assert "No such file" in capsys.readouterr().err


def test_defaults_follow_the_reference_flags():
    args = cli.build_parser().parse_args(["render"])
    assert args.scene is None and args.out == "render.exr"
    assert args.spp is None and args.width is None and args.height is None and args.depth is None


@pytest.mark.gpu
def test_render_writes_the_film_averages_as_a_multilayer_exr(ctx, tmp_path):
    import cenote_amd as cn
    from cenote_amd import workloads
    from test_exr import channels

    out = tmp_path / "cornell.exr"
    run = subprocess.run([sys.executable, "-m", "cenote_amd.cli", "render", "cornell", "--spp", "3", "--width", "48",
                          "--height", "32", "--out", str(out)], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    assert "48×32, 3 spp" in run.stdout
    _, names, ch = channels(out)
    assert names == ["A", "B", "G", "R", "Z", "albedo.B", "albedo.G", "albedo.R", "normal.X", "normal.Y", "normal.Z"]
    scene = cn.Scene(ctx, workloads.cornell())
    r = cn.Renderer(ctx, max_bounces=workloads.CONFIGS["cornell"]["max_bounces"])
    film = cn.Film(ctx, 48, 32)
    r.accumulate(scene, film, count=3)
    beauty, albedo, normal, depth = film.averages()
    film.close()
    r.close()
    scene.close()
    assert np.array_equal(ch["R"], beauty[..., 0].astype(np.float16))
    assert np.array_equal(ch["albedo.G"], albedo[..., 1].astype(np.float16))
    assert np.array_equal(ch["Z"].view(np.uint32), depth.view(np.uint32))


def test_a_broken_scene_file_is_an_error_before_any_device_work(tmp_path, capsys):
    bad = tmp_path / "bad.ron"
    bad.write_text("(version: 7, ops: [])")
    assert cli.main(["render", str(bad)]) == 2
    assert "version 7" in capsys.readouterr().err


@pytest.mark.gpu
def test_render_reads_a_scene_file(tmp_path):
    from test_exr import channels

    scene = tmp_path / "tri.ron"
    scene.write_text('''(version: 1, ops: [
        Settings((name: "main", resolution: Some((40, 30)), spp: Some(2), max_bounces: Some(3))),
        Camera((name: "main", position: Some((0.3, 0.3, 2.0)), look_at: Some((0.3, 0.3, 0.0)))),
        Mesh((name: "tri", source: Some(Inline(positions: [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)],
                                               triangles: [(0, 1, 2)])))),
        Material((name: "lamp", emission_luminance: Some(2.0))),
        Instance((name: "t", mesh: Some("tri"), material: Some("lamp"))),
    ])''')
    out = tmp_path / "tri.exr"
    run = subprocess.run([sys.executable, "-m", "cenote_amd.cli", "render", str(scene), "--out", str(out)],
                         cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    assert "40×30, 2 spp" in run.stdout
    _, _, ch = channels(out)
    r = ch["R"].astype(np.float32)
    assert r.shape == (30, 40) and r.max() > 1.0 and r.min() == 0.0  # the emitter in front of a black sky
