"""A view's camera paths take the slot space in 8-row bands walked column by
column (band_pixel / band_index in the device records header): the mapping
must be a permutation of the view's pixels with band_index its inverse, for
any size, and every 64 consecutive slots of a full band must cover one 8 x 8
block of pixels."""

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SHIM = r"""
#include "cn_records.h"
extern "C" void shim_band(uint32_t width, uint32_t height, uint32_t* pixel, uint32_t* back) {
    for (uint32_t i = 0; i < width * height; ++i) {
        pixel[i] = cn::band_pixel(i, width, height);
        back[i] = cn::band_index(pixel[i], width, height);
    }
}
"""


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("slots")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    lib = d / "libshim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'cenote_amd' / 'csrc'}", str(src), "-o", str(lib)], check=True, capture_output=True)
    so = C.CDLL(str(lib))
    so.shim_band.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    return so


@pytest.mark.parametrize("width,height", [(1, 1), (7, 3), (13, 9), (64, 8), (100, 17), (256, 256), (1280, 720)])
def test_band_order_is_a_permutation_with_its_inverse(shim, width, height):
    n = width * height
    pixel = np.empty(n, np.uint32)
    back = np.empty(n, np.uint32)
    shim.shim_band(width, height, pixel.ctypes.data, back.ctypes.data)
    assert np.array_equal(np.sort(pixel), np.arange(n, dtype=np.uint32))
    assert np.array_equal(back, np.arange(n, dtype=np.uint32))
    full = (height // 8) * width * 8  # slots in full bands
    for start in range(0, full - 63, 64 * max(1, full // 6400)):
        if (start % (width * 8)) % 8 or (start % (width * 8)) + 64 > width * 8:
            continue
        block = pixel[start:start + 64]
        xs, ys = block % width, block // width
        assert xs.max() - xs.min() == 7 and ys.max() - ys.min() == 7
