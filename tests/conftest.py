"""Shared fixtures. GPU tests are marked ``gpu`` and call libpwk through its C ABI."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "pwdft-rs_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
SI_UPF = os.path.join(GOLDEN, "pseudo", "Si.upf")
C_UPF = os.path.join(GOLDEN, "pseudo", "C.upf")
FE_UPF = os.path.join(GOLDEN, "pseudo", "Fe.upf")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpwk kernels)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def pwk_mod():
    import pwk

    return pwk


@pytest.fixture(scope="session")
def gpu(pwk_mod):
    # torch (used by a few tests for device buffers) ships its own HIP runtime; it has to
    # initialise the device before libpwk's runtime does
    try:
        import torch

        torch.cuda.is_available()
    except ImportError:
        pass
    acc = pwk_mod.GpuAccelerator(0)
    yield acc
    acc.close()
