import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def pytest_collection_modifyitems(config, items):
    # GPU tests are selected with -m gpu.  The device check goes through the
    # library's own HIP runtime (torch is never initialised on the device: its
    # wheel carries a second HIP runtime).
    if not any("gpu" in item.keywords for item in items):
        return
    try:
        from piro_amd.core import device_count
        have_gpu = device_count() > 0
    except Exception:  # pragma: no cover
        have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device in this process")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
