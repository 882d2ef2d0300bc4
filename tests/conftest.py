import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU check, skipped unless ASTR_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("ASTR_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU check (set ASTR_SLOW=1)")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)
