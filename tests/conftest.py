import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP backend on an MI355X (gfx950)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(O.ORACLE_SO):
        O.build()
    return O.Oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle as O
    if not O.Reference.available():
        pytest.skip("oracle/_ref (the compiled reference core) is not built here")
    return O.Reference()


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
