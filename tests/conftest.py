import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "yao-rs_amd"))
sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def seeded_state(nbits, seed):
    """complex128 state with re, im i.i.d. N(0,1) from `seed`, L2-normalised."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(1 << nbits) + 1j * rng.standard_normal(1 << nbits)
    return z / np.linalg.norm(z)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="session")
def gpu():
    import yaohip
    yaohip._abi.init()
    return yaohip
