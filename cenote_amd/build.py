"""Builds libcenote_hip.so (gfx950) in-tree with the Makefile beside the
sources. ``python -m cenote_amd.build``."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent


def build(jobs: int = 8, force: bool = True) -> Path:
    csrc = PACKAGE / "csrc"
    env = dict(os.environ)
    if force:
        subprocess.run(["make", "-C", str(csrc), "clean"], check=True, env=env,
                       stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", str(csrc), f"-j{min(jobs, 16)}"], check=True, env=env)
    out = PACKAGE / "libcenote_hip.so"
    if not out.exists():
        raise RuntimeError(f"build did not produce {out}")
    return out


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
