import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "embree-aarch64_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size parity case")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_probe():
    return load_golden("reference_probe.json")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def rtc():
    import rtcore
    rtcore.lib()
    return rtcore


@pytest.fixture(scope="session")
def device(rtc):
    d = rtc.Device("")
    yield d
    d.release()


def records_equal(a, b):
    """Bytewise equality of two record arrays (NaN-safe, sign-of-zero exact)."""
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def mismatches(a, b):
    ab = a.view(np.uint8).reshape(len(a), -1)
    bb = b.view(np.uint8).reshape(len(b), -1)
    return np.nonzero((ab != bb).any(axis=1))[0]
