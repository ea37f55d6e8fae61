import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


REFERENCE = "/root/reference"
STATE_DAT = os.path.join(REFERENCE, "koodit/GrossPitaevskii/state.dat")


@pytest.fixture(scope="session")
def state_dat():
    if not os.path.exists(STATE_DAT):
        pytest.skip("reference tree (state.dat) not present on this machine")
    return STATE_DAT
