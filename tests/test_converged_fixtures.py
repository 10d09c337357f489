"""The converged-image fixtures are the oracle's: a few of their pixels,
recomputed here at the full 4096 samples, match bit for bit, and the tiles
cover what the fixture says (the whole Cornell frame; at least 1% of the
chart and heavy frames)."""

import os

import numpy as np
import pytest

from cenote_amd import workloads
from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,share", [("cornell", 1.0), ("chart", 0.01), ("heavy", 0.01)])
def test_fixture_is_the_oracle_average(name, share):
    f = np.load(os.path.join(GOLDEN, f"converged_{name}.npz"))
    width, height, depth, spp = (int(f[k]) for k in ("width", "height", "max_bounces", "spp"))
    pixels = f["pixels"]
    assert spp == 4096
    assert len(pixels) >= share * width * height
    assert np.all(np.diff(pixels.astype(np.int64)) > 0) and int(pixels[-1]) < width * height
    pick = np.random.default_rng(7).choice(len(pixels), size=3, replace=False)
    sums = orc.OracleScene(workloads.build(name)).accumulate_pixels(width, height, pixels[pick], 0, spp,
                                                                    max_bounces=depth)
    assert np.array_equal(sums[:, :3] / np.float32(spp), f["mean"][pick])


@pytest.mark.parametrize("name", ["chart", "heavy"])
def test_tiles_cover_every_material_class_in_view(name):
    f = np.load(os.path.join(GOLDEN, f"converged_{name}.npz"))
    classes = set(str(c) for c in f["classes"])
    expected = {"chart": {"metal", "dielectric", "sky"}, "heavy": {"glass", "coat", "fuzz", "emitter", "sky"}}[name]
    assert expected <= classes
