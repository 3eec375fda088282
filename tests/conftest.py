import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running validation at full size")


def _gpu_count():
    try:
        from mfree_iwf_amd import device_count
        return device_count()
    except Exception:
        return 0


def pytest_collection_modifyitems(config, items):
    if any("gpu" in it.keywords for it in items) and _gpu_count() == 0:
        skip = pytest.mark.skip(reason="no gfx950 device visible")
        for it in items:
            if "gpu" in it.keywords:
                it.add_marker(skip)
