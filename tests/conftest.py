import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_runtest_setup(item):
    # Name the case before any GPU work starts, so a fault still leaves a trace.
    if "gpu" in item.keywords:
        print(f"\n[gpu-test] {item.nodeid}", flush=True)
