"""Pins for the oracle's environment tables (environment.rs:130-218): the
reference's tests (environment.rs:258-337) restated, plus what the
mathematics fixes — CDF inversion reproduces the pdf table, and a constant
sky's power telescopes to 4*pi*L."""

import numpy as np

from oracle import oracle as orc


def image(w, h, fn):
    t = np.ones((h, w, 4), np.float32)
    for row in range(h):
        for col in range(w):
            t[row, col, :3] = fn(col, row)
    return t


def test_tables_are_normalized_distributions():
    env = image(8, 4, lambda c, r: (1 + c + r * 8) * 0.1)
    m, cdf, pdfs, _ = orc.env_tables(env)
    assert m[0] == 0.0 and m[-1] == 1.0 and np.all(np.diff(m) >= 0)
    cdf = cdf.reshape(4, 9)
    assert np.all(cdf[:, 0] == 0.0) and np.all(cdf[:, 8] == 1.0)
    assert np.all(np.diff(cdf, axis=1) >= 0)
    assert abs(pdfs.astype(np.float64).mean() - 1.0) < 1e-6


def test_support_covers_the_bilinear_footprint():
    env = image(8, 4, lambda c, r: float(c == 0 and r == 0))
    _, _, pdfs, _ = orc.env_tables(env)
    pdfs = pdfs.reshape(4, 8)
    for row in range(4):
        for col in range(8):
            in_footprint = row <= 1 and col in (7, 0, 1)
            assert (pdfs[row, col] > 0) == in_footprint, (col, row)


def test_black_environment_has_no_mass():
    m, _, pdfs, power = orc.env_tables(np.array([[[0, 0, 0, 1]]], np.float32))
    assert power == 0.0
    assert np.all(pdfs == 0.0)
    assert m[-1] == 1.0


def test_constant_environment_power_is_exact():
    lum = 0.5 * (0.2722287 + 0.6740818 + 0.0536895)
    lib = orc.lib()
    import ctypes as C

    rgb = (C.c_float * 3)(0.5, 0.5, 0.5)
    lum = lib.orc_luminance(rgb)
    for w, h in [(1, 1), (8, 4), (13, 7)]:
        _, _, _, power = orc.env_tables(image(w, h, lambda c, r: 0.5))
        assert abs(power - 4 * np.pi * lum) < 1e-9, (w, h)


def test_weights_follow_the_reference_definition():
    """The pdf table is the 3x3-max-dilated luminance times sin(theta) at row
    centers, normalised to mean 1 (environment.rs:140-198), computed here
    independently in f64 from the texels."""
    rng = np.random.default_rng(3)
    h, w = 16, 32
    env = np.ones((h, w, 4), np.float32)
    env[..., :3] = rng.exponential(1.0, size=(h, w, 3)).astype(np.float32)
    lum = (env[..., 0] * np.float32(0.2722287) + env[..., 1] * np.float32(0.6740818)) + env[..., 2] * np.float32(0.0536895)
    lum = lum.astype(np.float64)
    padded = np.pad(lum, ((1, 1), (0, 0)), mode="edge")
    padded = np.concatenate([padded[:, -1:], padded, padded[:, :1]], axis=1)
    dil = np.max([padded[1 + dr:1 + dr + h, 1 + dc:1 + dc + w] for dr in (-1, 0, 1) for dc in (-1, 0, 1)], axis=0)
    weight = dil * np.sin((np.arange(h) + 0.5) / h * np.pi)[:, None]
    want = weight * (w * h) / weight.sum()
    _, _, pdfs, _ = orc.env_tables(env)
    assert np.allclose(pdfs.reshape(h, w), want, rtol=1e-6)


def test_cdf_inversion_reproduces_the_pdf_table():
    """Rows are drawn in proportion to row mass and columns to texel weight:
    stepping the CDFs at a fine uniform grid lands in each texel with
    frequency ~ pdf / (w*h)."""
    rng = np.random.default_rng(5)
    h, w = 8, 16
    env = np.ones((h, w, 4), np.float32)
    env[..., :3] = rng.uniform(0.0, 4.0, size=(h, w, 3)).astype(np.float32)
    m, cdf, pdfs, _ = orc.env_tables(env)
    cdf = cdf.reshape(h, w + 1)
    us = (np.arange(4000) + 0.5) / 4000
    rows = np.searchsorted(m, us, side="right") - 1
    row_freq = np.bincount(rows, minlength=h) / len(us)
    row_mass = pdfs.reshape(h, w).sum(axis=1) / pdfs.sum()
    assert np.allclose(row_freq, row_mass, atol=2e-3)
    for r in range(h):
        cols = np.searchsorted(cdf[r], us, side="right") - 1
        col_freq = np.bincount(np.minimum(cols, w - 1), minlength=w) / len(us)
        col_mass = pdfs.reshape(h, w)[r] / pdfs.reshape(h, w)[r].sum()
        assert np.allclose(col_freq, col_mass, atol=2e-3)
