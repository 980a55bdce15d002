import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libcvf.so)")
    config.addinivalue_line("markers", "slow: long CPU run, excluded from the default CPU suite")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("CVF_RUN_SLOW"):
        return
    skip = pytest.mark.skip(reason="slow CPU statistics run; set CVF_RUN_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)
