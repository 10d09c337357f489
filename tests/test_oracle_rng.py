"""Pins for the oracle's Sobol-Burley sampler (shaders/rng.slang).

Pins: the reference's own generator-matrix table (rng.slang:55-64), its
stratification test (wavefront.rs:1041-1136), the textbook values of the
Sobol sequence's first two dimensions, and bijectivity of the scramble —
so a wrong constant, index order or bit reversal fails one of them."""

import numpy as np

from oracle import oracle as orc

# rng.slang:55-64, the reference's literal table.
REFERENCE_SOBOL_DIM1 = [
    0x80000000, 0xc0000000, 0xa0000000, 0xf0000000, 0x88000000, 0xcc000000, 0xaa000000, 0xff000000,
    0x80800000, 0xc0c00000, 0xa0a00000, 0xf0f00000, 0x88880000, 0xcccc0000, 0xaaaa0000, 0xffff0000,
    0x80008000, 0xc000c000, 0xa000a000, 0xf000f000, 0x88008800, 0xcc00cc00, 0xaa00aa00, 0xff00ff00,
    0x80808080, 0xc0c0c0c0, 0xa0a0a0a0, 0xf0f0f0f0, 0x88888888, 0xcccccccc, 0xaaaaaaaa, 0xffffffff,
]


def test_generator_matrix_matches_the_reference_table():
    lib = orc.lib()
    for bit, column in enumerate(REFERENCE_SOBOL_DIM1):
        assert lib.orc_sobol(1 << bit, 1) == column


def test_first_dimensions_are_the_sobol_points():
    """The first 2^k unscrambled points are the 2-D Sobol point set: scipy's
    generator (Joe & Kuo direction numbers, Gray-code order) yields the same
    set, and dimension 0 is van der Corput."""
    from scipy.stats import qmc

    lib = orc.lib()
    n = 1024
    ours = {(lib.orc_sobol(i, 0) / 2**32, lib.orc_sobol(i, 1) / 2**32) for i in range(n)}
    theirs = {tuple(p) for p in qmc.Sobol(d=2, scramble=False).random(n).tolist()}
    assert ours == theirs
    want0 = [0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]
    for i, w in enumerate(want0):
        assert lib.orc_sobol(i, 0) / 2**32 == w


def test_reverse_bits_is_an_involution():
    lib = orc.lib()
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 2**32, size=200, dtype=np.uint64):
        x = int(x)
        assert lib.orc_reverse_bits(lib.orc_reverse_bits(x)) == x
    assert lib.orc_reverse_bits(1) == 0x80000000


def test_laine_karras_is_a_bijection_on_low_bits():
    """Each output bit depends only on the bits below it, so the map permutes
    every residue class mod 2^k."""
    lib = orc.lib()
    for seed in (0, 0x1234567, 0xdeadbeef):
        outs = {lib.orc_laine_karras(x, seed) & 0xfff for x in range(4096)}
        assert len(outs) == 4096


def test_sampler_is_stratified_and_decorrelated():
    """wavefront.rs:1041-1136 on the oracle: the first 64 points of three keys
    fill every 8x8 cell and every 1/64 bin of each axis exactly once."""
    for pixel, dimension in [(0, 0), (7, 0), (123456, 3)]:
        pts = np.array([orc.sample_2d(pixel, i, dimension) for i in range(64)])
        vals = np.array([orc.sample_1d(pixel, i, dimension) for i in range(64)])
        assert np.all((pts >= 0) & (pts < 1)) and np.all((vals >= 0) & (vals < 1))
        cells = np.bincount((pts[:, 1] * 8).astype(int) * 8 + (pts[:, 0] * 8).astype(int), minlength=64)
        assert np.all(cells == 1)
        for axis in (pts[:, 0], pts[:, 1], vals):
            assert np.all(np.bincount((axis * 64).astype(int), minlength=64) == 1)
    a = [orc.sample_2d(0, i, 0) for i in range(64)]
    assert a != [orc.sample_2d(1, i, 0) for i in range(64)]
    assert a != [orc.sample_2d(0, i, 1) for i in range(64)]


def test_unit_floats_are_24_bit_exact():
    """toUnitFloat (rng.slang:114-116): multiples of 2^-24 in [0, 1)."""
    for i in range(256):
        x, y = orc.sample_2d(99, i, 5)
        for v in (x, y):
            assert 0.0 <= v < 1.0
            assert float(v) * 2**24 == int(float(v) * 2**24)


def test_hash_uniform_range_and_independence():
    """rng.slang:128-132 keys (pixel, sample, bounce << 8 | crossing); the
    crossing cap (intersect.slang:23) keeps crossings under 256."""
    lib = orc.lib()
    vals = np.array([lib.orc_hash_uniform(3, s, b, c) for s in range(32) for b in range(16) for c in range(8)])
    assert np.all((vals >= 0) & (vals < 1))
    assert abs(vals.mean() - 0.5) < 0.02
    assert len(np.unique(vals)) > 4080
