import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through libxenon_hip.so")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    return pyoracle.Lib("strict")


@pytest.fixture(scope="session")
def reference():
    import pyoracle

    if not os.path.exists(pyoracle.PATHS["ref"]):
        pytest.skip("oracle/_ref is not built here (the reference sources are absent)")
    return pyoracle.Lib("ref")
