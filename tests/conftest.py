import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def orc():
    import subprocess
    from oracle import ORACLE_LIB, load_oracle
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return load_oracle()


def golden_case(golden, prefix):
    p = prefix + "/"
    return {k[len(p):]: v for k, v in golden.items() if k.startswith(p)}


@pytest.fixture(scope="session")
def hip():
    """libnqs_hip.so on a GPU machine; fails loudly when it is not built."""
    from spin_based_neural_network_amd import load_library
    return load_library()
