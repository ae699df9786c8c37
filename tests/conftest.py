"""Shared test set-up.

* ``gpu`` marks tests that need an MI355X; ``-m "not gpu"`` runs the rest on CPU.
* The oracle (oracle/) is the checker; the product under test is libcns_hip.so.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and libcns_hip.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.reference_available():
        pytest.skip("oracle/_ref not built (the reference tree is absent here)")
    return oracle.Reference()


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from cns_nospec_amd import _lib
    _lib.load()  # raises if the HIP library is not built: there is no fallback
    return torch.device("cuda:0")
