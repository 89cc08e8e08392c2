import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the HIP library")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


def pytest_runtest_logstart(nodeid, location):
    # name the case before it touches the GPU, so a failing run says which one
    print("\n[case] " + nodeid, flush=True)
