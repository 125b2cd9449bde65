import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(REPO, "tests", "golden", "mtgp_ref_words.npz"))


@pytest.fixture(scope="session")
def mtgp_table():
    import oracle
    return oracle.mtgp_params()


@pytest.fixture(scope="session")
def engine():
    import esthera_amd as ea
    e = ea.Engine(0)
    yield e
    e.close()
