"""The corpus through the HIP path (tests/corpus.rs:87-117): each vendored
pbrt scene imported, rendered at the golden's size to 64 spp by the
library, and compared with the reference's own golden render. The
reference's bar is mean FLIP <= 0.01; the same sampler gives the same
noise here, so the bar is per pixel: all but a few pixels agree to 1e-3."""

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cenote_amd as cn
from cenote_amd import pbrt, scenefile
from cenote_amd.output import read_exr

from test_pbrt import CORPUS, CORPUS_SCENES, corpus_render_error

pytestmark = pytest.mark.gpu


def imported(name):
    ops, _ = pbrt.import_scene(CORPUS / name / "scene-v4.pbrt")
    desc = scenefile.Description()
    desc.apply(ops)
    return scenefile.lower(desc)


@pytest.mark.parametrize("name,close,mean_abs", [("cornell-box", 0.999, 1e-5), ("veach-mis", 0.99, 5e-4)])
def test_corpus_renders_match_the_reference_goldens(ctx, name, close, mean_abs):
    data, settings = imported(name)
    w, h = CORPUS_SCENES[name]
    scene = cn.Scene(ctx, data)
    r = cn.Renderer(ctx, max_bounces=settings["max_bounces"])
    film = cn.Film(ctx, w, h)
    r.accumulate(scene, film, count=64)
    beauty = r.resolve(film)[0]
    film.close()
    r.close()
    scene.close()
    frac, mean = corpus_render_error(name, beauty)
    assert frac >= close and mean <= mean_abs, (frac, mean)


def test_the_cli_renders_a_pbrt_scene(tmp_path):
    out = tmp_path / "cornell.exr"
    proc = subprocess.run([sys.executable, "-m", "cenote_amd.cli", "render",
                           str(CORPUS / "cornell-box" / "scene-v4.pbrt"), "--width", "256", "--height", "256",
                           "--spp", "64", "--out", str(out)], cwd=Path(__file__).resolve().parents[1],
                          capture_output=True, text=True, timeout=110)
    assert proc.returncode == 0, proc.stderr
    # The first layer is the beauty, stored as f16 (at most 2^-11 relative).
    frac, _ = corpus_render_error("cornell-box", read_exr(out))
    assert frac >= 0.999
