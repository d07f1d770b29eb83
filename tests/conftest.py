import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_PATH = os.path.join(ROOT, "tests", "golden", "golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN_PATH) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def raja():
    """The HIP library on cuda:0; fails loudly when it cannot run."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a GPU")
    import raja_amd
    raja_amd.lib()
    return raja_amd


@pytest.fixture(autouse=True)
def _lookback_waits_completed(request):
    """After every test that used the GPU library: no scan or radix pass gave
    up waiting on a lookback descriptor (such a wait is bounded and recorded
    instead of hanging the device)."""
    yield
    if "raja" in request.fixturenames:
        request.getfixturevalue("raja").device_status()
