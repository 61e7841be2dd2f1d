"""Shared test setup: markers, repository path, GPU detection."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def _gpu_present() -> bool:
    if os.environ.get("PSCBFM_FORCE_NO_GPU"):
        return False
    return Path("/dev/kfd").exists()


GPU_PRESENT = _gpu_present()


def pytest_collection_modifyitems(config, items):
    if GPU_PRESENT:
        return
    skip = pytest.mark.skip(reason="no GPU on this machine")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
