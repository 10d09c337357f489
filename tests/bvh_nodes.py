"""Helpers for the BVH node tests: full-precision node records, the
library's quantizer, and an exact decoding of the quantized planes."""

import ctypes as C

import numpy as np

from cenote_amd import _abi

EMPTY = 0xFFFFFFFF


def full_nodes(boxes_lo, boxes_hi, children):
    """(N, 4, 3) low / high corners and (N, 4) child refs as 128-byte records."""
    n = len(children)
    rec = np.zeros((n, 32), np.uint32)
    f = rec.view(np.float32)
    for k in range(3):
        f[:, 8 * k: 8 * k + 4] = boxes_lo[:, :, k]
        f[:, 8 * k + 4: 8 * k + 8] = boxes_hi[:, :, k]
    rec[:, 24:28] = children
    rec[:, 28] = np.arange(n, dtype=np.uint32) % 40  # depth
    return rec


def quantize(rec):
    lib = _abi.library()
    out = np.zeros((len(rec), 16), np.uint32)
    _abi.check(lib.cn_bvh_quantize(rec.ctypes.data_as(C.c_void_p), len(rec), out.ctypes.data_as(C.c_void_p)))
    return out


def decode(q):
    """Boxes, grid steps, the steps' biased exponents and child refs of
    64-byte nodes: origin and x step in words 0..3, the byte planes in
    words 4..9, the y and z steps in words 10..11, the refs in 12..15."""
    origin = q[:, 0:3].view(np.float32).astype(np.float64)
    step = q[:, [3, 10, 11]].view(np.float32).astype(np.float64)
    mantissa, exponent = np.frexp(step)
    assert np.all(mantissa == 0.5), "every grid step is a power of two"
    exps = (exponent - 1 + 127).astype(np.int64)
    planes = q[:, [4, 5, 6, 7, 8, 9]]  # lo_x hi_x lo_y hi_y lo_z hi_z
    byte = np.stack([(planes >> (8 * c)) & 0xFF for c in range(4)], axis=2).astype(np.float64)  # (N, 6, 4)
    lo = np.stack([origin[:, k, None] + byte[:, 2 * k] * step[:, k, None] for k in range(3)], axis=2)
    hi = np.stack([origin[:, k, None] + byte[:, 2 * k + 1] * step[:, k, None] for k in range(3)], axis=2)
    return lo, hi, step, exps, q[:, 12:16]




def contains(full, q):
    """Whether every used child box of the 128-byte records `full` lies
    inside its quantized box in `q` (exact arithmetic)."""
    f = full.view(np.float32)
    lo = np.stack([f[:, 8 * k: 8 * k + 4] for k in range(3)], axis=2).astype(np.float64)
    hi = np.stack([f[:, 8 * k + 4: 8 * k + 8] for k in range(3)], axis=2).astype(np.float64)
    used = full[:, 24:28] != EMPTY
    qlo, qhi, _, _, qchild = decode(q)
    return bool(np.all((qlo <= lo)[used]) and np.all((qhi >= hi)[used]) and np.array_equal(qchild, full[:, 24:28]))
