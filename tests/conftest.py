"""Shared test setup: import paths for the package binding and the oracle, the `gpu` marker, fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "aerial-mimo-tc-optimization_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def po():
    import pe_oracle
    pe_oracle.lib()
    return pe_oracle


@pytest.fixture(scope="session")
def pe():
    import pusch_eq
    pusch_eq.lib()
    return pusch_eq


@pytest.fixture(scope="session")
def dev(pe):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    pe.init(0)
    return torch.device("cuda:0")
