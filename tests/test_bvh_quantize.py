"""The traversal's 64-byte BVH nodes (cn_bvh_quantize, host code): every
child box, quantized to bytes on its node's power-of-two grid, must contain
the full-precision box it stands for — the traversal is conservative only
if no plane moves inward. The checks decode the planes in exact (float64)
arithmetic from the fields alone."""

import numpy as np
import pytest

from bvh_nodes import EMPTY, decode, full_nodes, quantize


def random_case(rng, n, scale, offset):
    center = rng.uniform(-1, 1, size=(n, 1, 3)) * scale + offset
    size = np.exp(rng.uniform(-12, 0, size=(n, 4, 3))) * scale
    lo = (center + rng.uniform(-1, 1, size=(n, 4, 3)) * scale * 0.5).astype(np.float32)
    hi = (lo + size).astype(np.float32)
    children = rng.integers(0, 1 << 20, size=(n, 4), dtype=np.uint32)
    children[rng.random((n, 4)) < 0.2] = EMPTY
    children[:, 0] = rng.integers(0, 1 << 20, size=n, dtype=np.uint32)  # at least one child
    return lo, hi, children


@pytest.mark.parametrize("scale,offset", [(1.0, 0.0), (50.0, 0.0), (3e4, 1e5), (1e-3, -7.0)])
def test_quantized_boxes_contain_the_full_precision_boxes(scale, offset):
    rng = np.random.default_rng(int(scale * 1000) % 2**32)
    lo, hi, children = random_case(rng, 4000, scale, offset)
    q = quantize(full_nodes(lo, hi, children))
    qlo, qhi, step, exps, qchild = decode(q)
    assert np.array_equal(qchild, children)
    used = children != EMPTY
    assert np.all((qlo <= lo.astype(np.float64))[used])
    assert np.all((qhi >= hi.astype(np.float64))[used])
    # Outward by less than one grid step, and the grid is the finest that
    # spans the node in 255 steps.
    assert np.all((lo - qlo < step[:, None, :])[used])
    assert np.all((qhi - hi < step[:, None, :])[used])
    big = np.where(used[..., None], hi.astype(np.float64), -np.inf).max(axis=1)
    small = np.where(used[..., None], lo.astype(np.float64), np.inf).min(axis=1)
    extent = big - small
    assert np.all(255.0 * step >= extent)
    assert np.all((127.5 * step < extent) | (exps == 1))
    assert np.all(step >= 2.0 ** -126)  # normal floats: the traversal scales by them exactly


def test_flat_and_point_children_stay_inside_their_grid():
    lo = np.zeros((3, 4, 3), np.float32)
    hi = np.zeros((3, 4, 3), np.float32)
    hi[0, :, :2] = 1.0           # a flat node (zero extent in z)
    lo[1] = hi[1] = 2.5          # four coincident points
    lo[2, :, :] = [[-1, -1, -1], [0, 0, 0], [1, 1, 1], [2, 2, 2]]
    hi[2] = lo[2] + 0.25
    children = np.array([[1, 2, 3, 4], [5, EMPTY, EMPTY, EMPTY], [6, 7, 8, 9]], np.uint32)
    qlo, qhi, step, exps, _ = decode(quantize(full_nodes(lo, hi, children)))
    used = children != EMPTY
    assert np.all((qlo <= lo)[used]) and np.all((qhi >= hi)[used])
    assert exps[0, 2] == 1 and exps[1].tolist() == [1, 1, 1]  # zero extent: the smallest normal step
