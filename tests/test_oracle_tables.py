"""Pins for the BSDF table blob the oracle and the HIP path both read
(crates/cenote/src/tables/{ggx_energy,ltc_sheen}.bin, copied as data): the
reference's provenance and physicality tests (tables.rs:414-533), plus an
independent re-integration of GGX directional albedo (Fresnel = 1), the
quantity table GGX_E holds."""

import numpy as np

from oracle import oracle as orc

T = orc.bsdf_tables()
GGX_E, GGX_E_AVG, SHEEN_A, SHEEN_B, SHEEN_R = 0, 1024, 13856, 14880, 15904


def test_blob_sizes():
    assert T.size == 16928  # tables.rs:67-80


def test_sheen_table_matches_its_provenance():
    corners = [(0, 0, 0.014150003, 0.0006043148, 1.4e-5), (1, 0, 0.019405415, -0.002323515, 0.05839),
               (30, 31, 0.8667744, 1.8694525e-5, 0.315084), (31, 31, 0.8795804, 2.7911458e-5, 0.341875)]
    for rough, mu, a, b, r in corners:
        i = rough * 32 + mu
        assert abs(T[SHEEN_A + i] - a) < 1e-6
        assert abs(T[SHEEN_B + i] - b) < 1e-6
        assert abs(T[SHEEN_R + i] - r) < 1e-6
    degenerate = np.abs(T[SHEEN_A:SHEEN_A + 1024]) < 1e-5
    assert np.all(T[SHEEN_R:SHEEN_R + 1024][degenerate] < 1e-3)


def test_energy_tables_are_physical():
    e = T[:13856]
    assert np.all((e >= 0) & (e <= 1 + 1e-3))
    ggx = T[GGX_E:GGX_E + 1024].reshape(32, 32)  # [mu][rough]
    assert np.allclose(ggx[:, 0], 1.0, atol=2e-3)
    for mu in (8, 31):
        assert ggx[mu, 4] > ggx[mu, 31]


def _ggx_albedo(rough, mu, n=256):
    """Directional albedo of D*G1*G1/(4 mu_o mu_i) with F = 1, by VNDF
    importance sampling on a stratified grid (weight = G1(wi))."""
    alpha = max(rough, 0.035) ** 2
    wo = np.array([np.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])
    u1, u2 = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n)
    u1, u2 = u1.ravel(), u2.ravel()
    # Heitz 2018 VNDF sampling
    vh = np.array([alpha * wo[0], alpha * wo[1], wo[2]])
    vh /= np.linalg.norm(vh)
    lensq = vh[0] ** 2 + vh[1] ** 2
    t1 = np.array([-vh[1], vh[0], 0]) / np.sqrt(lensq) if lensq > 0 else np.array([1.0, 0, 0])
    t2 = np.cross(vh, t1)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    p1 = r * np.cos(phi)
    p2 = r * np.sin(phi)
    s = 0.5 * (1 + vh[2])
    p2 = (1 - s) * np.sqrt(np.maximum(0, 1 - p1 * p1)) + s * p2
    nh = p1[:, None] * t1 + p2[:, None] * t2 + np.sqrt(np.maximum(0, 1 - p1 * p1 - p2 * p2))[:, None] * vh
    h = np.stack([alpha * nh[:, 0], alpha * nh[:, 1], np.maximum(1e-6, nh[:, 2])], axis=1)
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    wi = 2 * (h @ wo)[:, None] * h - wo
    a2 = alpha * alpha
    g1 = 2 * np.abs(wi[:, 2]) / (np.abs(wi[:, 2]) + np.sqrt(wi[:, 2] ** 2 * (1 - a2) + a2))
    return float(np.where(wi[:, 2] > 0, g1, 0.0).mean())


def test_ggx_energy_matches_an_independent_integration():
    """tables.rs:414-470's spot check, integrated here with a different
    sampler (Heitz's VNDF) — grid point k of 32 sits at k/31."""
    for mu_i, rough_i in [(31, 8), (16, 16), (4, 31), (8, 24)]:
        mu, rough = mu_i / 31, rough_i / 31
        want = _ggx_albedo(rough, mu)
        got = T[GGX_E + mu_i * 32 + rough_i]
        assert abs(got - want) < 3e-3, (mu_i, rough_i, got, want)
