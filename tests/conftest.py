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


def pytest_collection_modifyitems(config, items):
    # print the case name before each GPU test so a box that dies mid-run
    # still names the test in its last output line
    for it in items:
        if it.get_closest_marker("gpu"):
            it.add_marker(pytest.mark.usefixtures("_announce"))


@pytest.fixture
def _announce(request):
    print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
