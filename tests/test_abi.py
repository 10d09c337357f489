"""The C ABI without a GPU: the library loads, exports every entry point
include/cenote_hip.h declares, the record layouts match the reference's
byte sizes, and calls without a device fail with a status, not a crash."""

import ctypes as C
import re
from pathlib import Path

import pytest

from cenote_amd import _abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "cenote_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("cn_context_create", "cn_scene_create", "cn_renderer_create", "cn_trace", "cn_accumulate",
                     "cn_film_averages", "cn_resolve", "cn_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = _abi.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(declared_functions()) == set(_abi.SIGNATURES)


def test_record_sizes_match_the_reference():
    assert C.sizeof(_abi.Material) == 144      # material.rs:116
    assert C.sizeof(_abi.LightRecord) == 80    # lights.rs:115
    assert C.sizeof(_abi.Instance) == 4 + 4 + 48 + 144


def test_errors_are_statuses_not_crashes():
    lib = _abi.library()
    assert lib.cn_context_create(0, None) == _abi.CN_E_INVALID
    assert b"null" in lib.cn_last_error()
    assert lib.cn_scene_create(None, None, None) == _abi.CN_E_INVALID
    assert lib.cn_renderer_create(None, 0, 8, 0, None) == _abi.CN_E_INVALID


def test_context_without_a_device_reports_no_device():
    import os

    lib = _abi.library()
    handle = C.c_void_p()
    status = lib.cn_context_create(0, C.byref(handle))
    if status == _abi.CN_OK:
        pytest.skip("a device is visible here")
    assert status == _abi.CN_E_NO_DEVICE, lib.cn_last_error()
