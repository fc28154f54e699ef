import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def mf():
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from orc import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def gpu(mf):
    rc = mf.lib().mf_init(0)
    if rc != 0:
        raise RuntimeError("mf_init failed on a GPU test: " + mf.lib().mf_last_error().decode())
    return mf
