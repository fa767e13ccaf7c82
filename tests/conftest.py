"""Test configuration: the `gpu` marker, import paths, and shared fixtures."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "diplodocustransport.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_modifyitems(config, items):
    # GPU tests first announce themselves so a hang names its case
    for item in items:
        if item.get_closest_marker("gpu"):
            item.add_marker(pytest.mark.timeout(300))


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
