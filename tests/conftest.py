import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch ships its own HIP runtime (libamdhip64 / libhsa-runtime64 under
# torch/lib).  Importing it before libcgsense_hip.so is loaded makes the library
# bind to that same runtime (its libamdhip64.so.7 dependency is then already
# present), as in bench.py; loading the system runtime first and torch later
# would put two HIP runtimes in one process and torch finds no GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ora():
    from oracle import ora as _ora
    _ora.build()
    return _ora


@pytest.fixture(scope="session")
def cgs():
    import codeare_amd
    codeare_amd.load_library()
    return codeare_amd
