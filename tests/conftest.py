import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libhgx.so")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    O.lib()
    return O


@pytest.fixture(scope="session")
def hgx():
    # torch first: it ships its own HIP runtime, and a process that loads
    # libhgx.so first binds torch to the library's runtime instead, which
    # torch cannot use; loaded in this order both share torch's
    import torch  # noqa: F401

    import hypergeometrica_amd as h
    from hypergeometrica_amd import _lib
    _lib.lib()  # initialise device 0; raises if the library or device is missing
    return h
