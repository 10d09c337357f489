"""EXR output and input over the library (output.rs:37-193): linear ACEScg
files tagged with the AP1 chromaticities. Images are numpy arrays — (H, W,
4) RGBA f32 and (H, W) depth — with pixel (0, 0) top-left."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import check, library


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _path(path) -> bytes:
    return os.fsencode(os.fspath(path))


def write_exr(path, rgba: np.ndarray):
    """output.rs:37-57: RGBA f32 channels."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w, _ = rgba.shape
    check(library().cn_write_exr(_path(path), w, h, _fp(rgba)))


def write_aov_exr(path, beauty, albedo, normal, depth):
    """output.rs:84-136: f16 R/G/B/A, f32 Z, f16 albedo.R/G/B, normal.X/Y/Z."""
    beauty, albedo, normal = (np.ascontiguousarray(x, dtype=np.float32) for x in (beauty, albedo, normal))
    depth = np.ascontiguousarray(depth, dtype=np.float32)
    h, w, _ = beauty.shape
    for x in (albedo, normal):
        if x.shape != beauty.shape:
            raise ValueError("AOV layers must match the beauty's shape")
    if depth.shape != (h, w):
        raise ValueError("depth must be (height, width)")
    check(library().cn_write_aov_exr(_path(path), w, h, _fp(beauty), _fp(albedo), _fp(normal), _fp(depth)))


def read_exr(path) -> np.ndarray:
    """output.rs:148-193: the first layer's RGBA as f32 (missing alpha = 1)."""
    lib = library()
    w, h = C.c_uint32(), C.c_uint32()
    check(lib.cn_read_exr(_path(path), C.byref(w), C.byref(h), None, 0))
    out = np.zeros((h.value, w.value, 4), np.float32)
    check(lib.cn_read_exr(_path(path), C.byref(w), C.byref(h), _fp(out), out.size))
    return out
