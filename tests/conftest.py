"""pytest configuration: the `gpu` marker and import paths.

GPU tests call the HIP library through its C ABI; they are skipped only when
no GPU device exists at all (this container), never when the library is
missing -- then they fail.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("for-qaoa_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_HAS_GPU = None


def has_gpu():
    global _HAS_GPU
    if _HAS_GPU is None:
        # /dev/kfd is the ROCm compute device node; without it no HIP device exists
        _HAS_GPU = os.path.exists("/dev/kfd")
    return _HAS_GPU


def pytest_collection_modifyitems(config, items):
    if has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU device in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
