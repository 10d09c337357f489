"""Writes the converged-image fixtures that tests/test_gpu_converged.py
compares the device against: the oracle's 4096-sample beauty average over

- the whole Cornell frame at 256 x 256 (depth 8);
- seeded 8 x 8 tiles covering at least 1% of the chart frame (1280 x 720,
  depth 12) and of the heavy frame (1920 x 1080, depth 16), chosen so that
  every material class the bounce-0 hits show (glass, coated metal, fuzz,
  metal, dielectric, emitter, sky) has tiles of its own, the rest at random.

Each fixture holds the frame size, depth, sample count, the listed pixel
indices (row-major) and their averaged RGB. Run from the repository root:

    python tests/golden/make_converged.py [cornell chart heavy]

It takes several minutes of CPU (the oracle traces every listed pixel's
4096 paths on all host cores).
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cenote_amd import workloads  # noqa: E402
from oracle import oracle as orc  # noqa: E402

SPP = 4096
TILE = 8
SEED = 2024
# (share of pixels covered, tiles per material class)
TILED = {"chart": (0.0105, 20), "heavy": (0.0105, 40)}


def material_class(m) -> str:
    if max(m.emission) > 0.0:
        return "emitter"
    if m.transmission_weight > 0.0:
        return "glass"
    if m.coat_weight > 0.0:
        return "coat"
    if m.fuzz_weight > 0.0:
        return "fuzz"
    if m.metalness > 0.0:
        return "metal"
    return "dielectric"


def pixel_classes(data, width, height) -> np.ndarray:
    """The material class name of every pixel's bounce-0 hit (sample 0)."""
    inst, _, _ = orc.OracleScene(data).first_hits(width, height, 0)
    names = np.array([material_class(i.material) for i in data.instances] + ["sky"])
    idx = np.where(inst < len(data.instances), inst, len(data.instances))
    return names[idx]


def choose_tiles(classes: np.ndarray, share: float, per_class: int, seed: int) -> np.ndarray:
    """Seeded TILE x TILE tiles: up to ``per_class`` among the tiles that
    hold each class, then random ones until ``share`` of the frame is
    covered. Returns the sorted pixel indices of the chosen tiles."""
    height, width = classes.shape
    tw, th = width // TILE, height // TILE
    assert tw * TILE == width and th * TILE == height
    rng = np.random.default_rng(seed)
    per_tile = classes.reshape(th, TILE, tw, TILE).transpose(0, 2, 1, 3).reshape(th * tw, TILE * TILE)
    chosen: set = set()
    for name in sorted(set(classes.ravel())):
        holders = np.flatnonzero((per_tile == name).any(axis=1))
        pick = rng.choice(holders, size=min(per_class, len(holders)), replace=False)
        chosen.update(int(t) for t in pick)
    need = int(np.ceil(share * width * height / (TILE * TILE)))
    while len(chosen) < need:
        chosen.add(int(rng.integers(th * tw)))
    tiles = np.array(sorted(chosen))
    ty, tx = tiles // tw, tiles % tw
    dy, dx = np.meshgrid(np.arange(TILE), np.arange(TILE), indexing="ij")
    rows = (ty[:, None, None] * TILE + dy[None]).ravel()
    cols = (tx[:, None, None] * TILE + dx[None]).ravel()
    return np.sort((rows * width + cols).astype(np.uint32))


def fixture(name: str) -> dict:
    cfg = workloads.CONFIGS[name]
    width, height, depth = cfg["width"], cfg["height"], cfg["max_bounces"]
    data = workloads.build(name)
    if name in TILED:
        share, per_class = TILED[name]
        classes = pixel_classes(data, width, height)
        pixels = choose_tiles(classes, share, per_class, SEED)
        present = sorted(set(classes.ravel()[pixels]))
    else:
        pixels = np.arange(width * height, dtype=np.uint32)
        present = []
    sums = orc.OracleScene(data).accumulate_pixels(width, height, pixels, 0, SPP, max_bounces=depth)
    mean = (sums[:, :3] / np.float32(SPP)).astype(np.float32)
    return dict(width=width, height=height, max_bounces=depth, spp=SPP, pixels=pixels, mean=mean,
                classes=np.array(present))


def main(names):
    for name in names:
        t = time.time()
        f = fixture(name)
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"converged_{name}.npz")
        np.savez_compressed(path, **f)
        share = len(f["pixels"]) / (f["width"] * f["height"])
        print(f"{name}: {len(f['pixels'])} pixels ({share:.2%}), classes {list(f['classes'])}, "
              f"{time.time() - t:.0f} s -> {path}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["chart", "heavy", "cornell"])
