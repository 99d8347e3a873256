import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import _pkg  # noqa: E402

_pkg.load()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libsvx.so")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0.  GPU tests fail loudly if the library or the device is missing."""
    import torch
    from u55c_statevector_30q_amd import libsvx
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    caps = libsvx.init(0)
    assert caps.arch.decode().startswith("gfx950"), caps.arch
    return caps
