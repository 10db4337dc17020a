import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _paths  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libscft on the GPU)")


@pytest.fixture(scope="session")
def abi():
    import abi as _abi
    return _abi


@pytest.fixture(scope="session")
def configs():
    import configs as _c
    return _c


@pytest.fixture(scope="session")
def orc(abi):
    import oracle
    return oracle.Oracle(abi)


@pytest.fixture(scope="session")
def lib(abi):
    """libscft.so; GPU tests fail loudly when it is missing."""
    return abi.load_library()
