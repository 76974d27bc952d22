import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "tt-ldpc-qb2-sim_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"
for p in (PKG_DIR, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


@pytest.fixture(scope="session")
def code12():
    from ldpc_mi355x import codes
    return codes.ar4ja(1024, "1/2")


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(GOLDEN_DIR, "stepper_profile_t.npz"))


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


def channel_params(ebn0_db, rate=0.5):
    """mu, sigma as fp32 exactly as libldpc computes them (ldpc_sim_scalar.cpp:352-355)."""
    g = 10.0 ** (ebn0_db / 10.0)
    return np.float32(4.0 * rate * g), np.float32(np.sqrt(8.0 * rate * g))


@pytest.fixture(scope="session")
def engine():
    """A Profile T engine on cuda:0 (GPU tests only)."""
    from ldpc_mi355x import decoder
    eng = decoder.Engine(1024, "1/2", max_iter=200, beta=0.9375)
    yield eng
    eng.close()
