"""Every record the C ABI passes by pointer has the same size and field
offsets in C (include/cenote_hip.h compiled by the host compiler) as in the
ctypes mirror the Python host uses: a C probe prints sizeof/offsetof, and
each ctypes field must sit where the C compiler put it."""

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from cenote_amd import _abi

HEADER_DIR = Path(__file__).resolve().parents[1] / "include"

PAIRS = {
    "cn_material": _abi.Material,
    "cn_mesh": _abi.Mesh,
    "cn_instance": _abi.Instance,
    "cn_delta_light": _abi.DeltaLight,
    "cn_environment": _abi.Environment,
    "cn_camera": _abi.Camera,
    "cn_scene_desc": _abi.SceneDesc,
    "cn_light_record": _abi.LightRecord,
    "cn_scene_info": _abi.SceneInfo,
    "cn_wave_counts": _abi.WaveCounts,
    "cn_traversal_counts": _abi.TraversalCounts,
    "cn_kernel_times": _abi.KernelTimes,
    "cn_launch_shape": _abi.LaunchShape,
    "cn_view_job": _abi.ViewJob,
    "cn_request_desc": _abi.RequestDesc,
    "cn_request_status": _abi.RequestStatus,
    "cn_server_stats": _abi.ServerStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "cenote_hip.h"', "int main(void) {"]
    for cname, cls in PAIRS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = d / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "probe"
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", str(HEADER_DIR), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        struct, field, value = line.split()
        table[(struct, field)] = int(value)
    return table


@pytest.mark.parametrize("cname", sorted(PAIRS))
def test_ctypes_mirror_matches_the_c_layout(c_layout, cname):
    cls = PAIRS[cname]
    assert C.sizeof(cls) == c_layout[(cname, "sizeof")]
    for field, _ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[(cname, field)], field
