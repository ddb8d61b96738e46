import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def lib_built():
    from tw_amd import api
    if not api.LIB_PATH.exists():
        from tw_amd import build
        build.build()
    return api.lib()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def pytest_runtest_setup(item):
    # name each GPU test before any of its fixtures run, past pytest's output
    # capture, so a run that dies or stalls mid-test still names the test in
    # its last output line
    if item.get_closest_marker("gpu"):
        capman = item.config.pluginmanager.getplugin("capturemanager")
        with capman.global_and_fixture_disabled():
            print(f"\n[gpu-test] {item.nodeid}", flush=True)
