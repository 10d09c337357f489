"""Pins for the oracle's light table (crates/cenote/src/lights.rs:175-224):
the reference's own tests (lights.rs:320-475) restated, plus the alias
table's defining property checked exactly in f64."""

import numpy as np

from cenote_amd import _abi
from oracle import oracle as orc

X, Y, Z = np.eye(3, dtype=np.float32)
O = np.zeros(3, np.float32)
LIGHT_NONE = 0xFFFFFFFF


def tri(a, b, c):
    return np.concatenate([a, b, c]).astype(np.float32)


def fields(rec):
    f = rec.view(np.float32)
    return dict(anchor=f[0:3], pdf=f[3], edge1=f[4:7], threshold=f[7], edge2=f[8:11], alias=rec[11],
                emission=f[12:15], instance=rec[15], kind=rec[16], primitive=rec[17])


def selection_probability(recs, index):
    """lights.rs:298-314: each slot gives `threshold` to itself and the rest to its alias."""
    mass = 0.0
    for slot, r in enumerate(recs):
        f = fields(r)
        t = float(f["threshold"])
        mass += (t if slot == index else 0.0) + ((1.0 - t) if f["alias"] == index else 0.0)
    return mass / len(recs)


def power(rec):
    """lights.rs:123-130 in f64, from the record's own f32 fields."""
    f = fields(rec)
    luma = float(np.float32(np.float32(f["emission"][0] * np.float32(0.2722287)) +
                            np.float32(f["emission"][1] * np.float32(0.6740818))) +
                 np.float32(f["emission"][2] * np.float32(0.0536895)))
    if f["kind"] == 0:
        area = float(np.linalg.norm(np.cross(f["edge1"].astype(np.float64), f["edge2"].astype(np.float64)))) / 2
        return np.pi * luma * area
    if f["kind"] == 1:
        return luma
    return 4 * np.pi * luma


def test_selection_matches_power_share():
    corners = [tri(O, X, Z), tri(O, [1, 0.5, 0], Z * 2), tri(Y, Y + X * 3, Y + Z)]
    emission = [[4, 4, 4], [1, 1, 1], [2, 0.5, 3]]
    ids = [[0, 0], [0, 1], [3, 0]]
    deltas = [(_abi.LIGHT_DISTANT, (0, -1, 0), (2, 2, 2)), (_abi.LIGHT_POINT, (0, 5, 0), (0.25, 0.25, 0.25))]
    recs, total = orc.build_lights(np.array(corners), np.array(emission), np.array(ids), deltas)
    assert len(recs) == 5
    powers = np.array([power(r) for r in recs])
    assert abs(powers.sum() - total) / total < 1e-6
    for i in range(5):
        assert abs(selection_probability(recs, i) - powers[i] / powers.sum()) < 1e-6


def test_pdf_folds_selection_per_kind():
    recs, _ = orc.build_lights(np.array([tri(O, X * 2, Z), tri(O, X * 6, Z)]), np.ones((2, 3)),
                               np.array([[0, 0], [7, 0]]))
    assert abs(fields(recs[0])["pdf"] - 0.25) < 1e-6
    assert abs(fields(recs[1])["pdf"] - 0.25) < 1e-6
    assert fields(recs[1])["instance"] == 7
    recs, _ = orc.build_lights(np.array([tri(O, X * 2, Z)]), np.ones((1, 3)), np.array([[0, 0]]),
                               [(_abi.LIGHT_POINT, (0, 1, 0), (0.25, 0.25, 0.25))])
    assert abs(fields(recs[1])["pdf"] - 0.5) < 1e-6
    assert fields(recs[1])["instance"] == LIGHT_NONE


def test_triangle_records_keep_input_order():
    recs, _ = orc.build_lights(np.array([tri(O, X, Z)] * 3), np.ones((3, 3)), np.array([[2, 0], [2, 1], [5, 0]]),
                               [(_abi.LIGHT_DISTANT, (0, -1, 0), (1, 1, 1))])
    ids = [(fields(r)["kind"], fields(r)["instance"], fields(r)["primitive"]) for r in recs]
    assert ids == [(0, 2, 0), (0, 2, 1), (0, 5, 0), (1, LIGHT_NONE, 0)]


def test_single_light_never_aliases():
    recs, _ = orc.build_lights(np.array([tri(O, X * 2, Z * 2)]), np.ones((1, 3)), np.array([[0, 0]]))
    assert abs(fields(recs[0])["threshold"] - 1.0) < 1e-7
    assert abs(fields(recs[0])["pdf"] - 0.5) < 1e-6


def test_degenerate_triangles_hold_their_slot_unsampled():
    recs, _ = orc.build_lights(np.array([tri(O, X, X * 2), tri(O, X, Z)]), np.ones((2, 3)),
                               np.array([[0, 0], [0, 1]]))
    assert fields(recs[0])["pdf"] == 0.0
    assert abs(selection_probability(recs, 0)) < 1e-12
    assert abs(selection_probability(recs, 1) - 1.0) < 1e-9


def test_a_powerless_list_builds_finite():
    recs, total = orc.build_lights(np.array([tri(O, X, X * 2)]), np.ones((1, 3)), np.array([[0, 0]]))
    assert fields(recs[0])["pdf"] == 0.0
    assert abs(fields(recs[0])["threshold"] - 1.0) < 1e-7
    assert abs(total) < 1e-12


def test_many_lights_alias_mass_is_exact():
    """10,000 random triangles (the heavy scene's scale): every light's alias
    mass equals its power share, and thresholds lie in [0, 1]."""
    rng = np.random.default_rng(7)
    n = 10_000
    corners = rng.uniform(-5, 5, size=(n, 9)).astype(np.float32)
    emission = np.repeat(np.exp(rng.uniform(np.log(1e-2), np.log(1e2), size=(n, 1))), 3, axis=1)
    ids = np.stack([np.arange(n), np.zeros(n)], axis=1)
    recs, total = orc.build_lights(corners, emission, ids)
    f = recs.view(np.float32)
    thresholds = f[:, 7].astype(np.float64)
    alias = recs[:, 11].astype(np.int64)
    assert np.all((thresholds >= 0) & (thresholds <= 1))
    mass = thresholds.copy()
    np.add.at(mass, alias, 1.0 - thresholds)
    powers = np.array([power(r) for r in recs])
    assert np.allclose(mass / n, powers / powers.sum(), atol=1e-6)


def test_select_probability_endpoints():
    """scene/mod.rs:665-671: exact 0 and 1 when the list or the sky is dark."""
    lib = orc.lib()
    assert lib.orc_select_probability(3.0, 0.0) == 1.0
    assert lib.orc_select_probability(0.0, 0.0) == 0.0
    assert lib.orc_select_probability(0.0, 2.0) == 0.0
    assert abs(lib.orc_select_probability(1.0, 3.0) - 0.25) < 1e-7
