import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_sessionstart(session):
    # torch reports no GPU if the HIP runtime was first initialised by the
    # library, so torch comes up first (a no-op without a GPU)
    import torch
    torch.cuda.is_available()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libpt_gravity.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a crash can be attributed
    if request.node.get_closest_marker("gpu"):
        sys.stdout.write("\n[case] %s\n" % request.node.nodeid)
        sys.stdout.flush()
    yield
