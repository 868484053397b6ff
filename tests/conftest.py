import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "sv-sim_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

os.environ.setdefault("SVSIM_QUIET", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs through libsvmi.so)")


def _gpu_present():
    # /dev/kfd is the ROCm compute device; no HIP call is made here.
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if _gpu_present():
        return
    skip = pytest.mark.skip(reason="no GPU in this container")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
