"""Test harness: GPU marker, import paths, oracle build.

CPU tests (``-m "not gpu"``) check the oracle against the reference's golden
vectors and the mathematics, the host-side mesh/RCM code and that the C-ABI
library exports its declared symbols.  GPU tests (``-m gpu``) compare the HIP
path, called through the C ABI, with the oracle on the same seeded inputs.
"""
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "zoomy-jax_amd"))
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.exists():
        subprocess.check_call(["make", "-s", "-C", str(ROOT / "oracle")])


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(GOLDEN / f"refs_{name}.npz")
    return load
