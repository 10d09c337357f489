"""Shared test setup: the repository root on sys.path, the `gpu` marker, and
session-scoped device handles for the GPU parity tests."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libcenote_hip.so")


@pytest.fixture(scope="session")
def ctx():
    import cenote_amd as cn

    context = cn.Context(0)
    yield context
    context.close()
