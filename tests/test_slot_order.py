"""A view's camera paths take the slot space in 8-row bands walked column by
column (band_pixel / band_index in the device records header): the mapping
must be a permutation of the view's pixels with band_index its inverse, for
any size, and every 64 consecutive slots of a full band must cover one 8 x 8
block of pixels. A job's samples then share the slots group by group
(job_slot / job_unslot): a bijection onto the job's slot span, the identity
with one sample, and each wave64 of slots a few pixels' consecutive
samples."""

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SHIM = r"""
#include "cn_records.h"
extern "C" void shim_job(uint32_t pixels, uint32_t samples, uint32_t* slot, uint32_t* index, uint32_t* k) {
    for (uint32_t i = 0; i < pixels; ++i)
        for (uint32_t s = 0; s < samples; ++s) slot[i * samples + s] = cn::job_slot(i, s, pixels, samples);
    for (uint32_t l = 0; l < pixels * samples; ++l) cn::job_unslot(l, pixels, samples, &index[l], &k[l]);
}
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
    so.shim_job.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
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


@pytest.mark.parametrize("pixels,samples", [(1, 1), (1, 5), (7, 3), (8, 8), (13, 1), (13, 32), (100, 17),
                                            (4096, 32), (1000, 64)])
def test_job_slots_are_a_bijection_with_their_inverse(shim, pixels, samples):
    n = pixels * samples
    slot = np.empty(n, np.uint32)
    index = np.empty(n, np.uint32)
    k = np.empty(n, np.uint32)
    shim.shim_job(pixels, samples, slot.ctypes.data, index.ctypes.data, k.ctypes.data)
    assert np.array_equal(np.sort(slot), np.arange(n, dtype=np.uint32))
    # job_unslot inverts job_slot
    want = slot.reshape(pixels, samples)
    assert np.array_equal(want[index, k], np.arange(n, dtype=np.uint32))
    if samples == 1:
        assert np.array_equal(slot, np.arange(n, dtype=np.uint32))
    # a pixel's samples sit at one stride, in sample order
    steps = np.diff(want.astype(np.int64), axis=1)
    assert np.all(steps > 0) and np.all(steps == steps[:, :1])
    # every wave64 of a full group span holds a handful of pixels
    if pixels >= 64 and samples >= 8:
        waves = index[: (n // 64) * 64].reshape(-1, 64)
        assert max(len(np.unique(w)) for w in waves) <= 16
