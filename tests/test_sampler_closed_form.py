"""The device sampler's Sobol axis 1 is a closed form (five shift/mask/xor
steps and a bit reversal) of the reference's column loop (rng.slang:100-111);
this compiles the device header's host path and checks it, and the whole
2-D draw built on it, against the oracle's loop on every low index and on
random 32-bit ones (the scramble makes indices use all 32 bits)."""

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = Path(__file__).resolve().parents[1]
SHIM = r"""
#include "cn_math.h"
extern "C" void shim_sobol_axis1(const uint32_t* in, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = cn::sobol_axis1(in[i]);
}
extern "C" void shim_sample_2d(uint32_t pixel, const uint32_t* samples, uint32_t dimension, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const cn::f2 f = cn::sample_2d(pixel, samples[i], dimension);
        out[2 * i] = f.x;
        out[2 * i + 1] = f.y;
    }
}
"""


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("sampler")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    lib = d / "libshim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", f"-I{ROOT / 'cenote_amd' / 'csrc'}", str(src), "-o", str(lib)],
                   check=True, capture_output=True)
    so = C.CDLL(str(lib))
    so.shim_sobol_axis1.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    so.shim_sample_2d.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]
    return so


def _indices():
    rng = np.random.default_rng(20261202)
    low = np.arange(1 << 12, dtype=np.uint32)
    wide = rng.integers(0, 1 << 32, size=20000, dtype=np.uint64).astype(np.uint32)
    edges = np.array([0xffffffff, 0x80000000, 0x7fffffff, 0xaaaaaaaa, 0x55555555], dtype=np.uint32)
    return np.concatenate([low, wide, edges])


def test_sobol_axis1_closed_form_matches_the_oracle_loop(shim):
    lib = orc.lib()
    idx = _indices()
    out = np.empty_like(idx)
    shim.shim_sobol_axis1(idx.ctypes.data, out.ctypes.data, idx.size)
    want = np.array([lib.orc_sobol(int(i), 1) for i in idx], dtype=np.uint32)
    assert np.array_equal(out, want)


def test_sample_2d_matches_the_oracle_bitwise(shim):
    lib = orc.lib()
    samples = np.arange(512, dtype=np.uint32)
    want = np.empty(2, dtype=np.float32)
    for pixel, dimension in ((0, 0), (7, 9), (123456, 21), (921599, 103)):
        got = np.empty(2 * samples.size, dtype=np.float32)
        shim.shim_sample_2d(pixel, samples.ctypes.data, dimension, got.ctypes.data, samples.size)
        for i, s in enumerate(samples):
            lib.orc_sample_2d(pixel, int(s), dimension, want.ctypes.data_as(C.POINTER(C.c_float)))
            assert got[2 * i] == want[0] and got[2 * i + 1] == want[1], (pixel, dimension, int(s))
