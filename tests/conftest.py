import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_fixture_arrays():
    """The reference fixture tests/random_10site_mps.dat as plain arrays."""
    d = np.load(os.path.join(GOLDEN, "random_10site_mps.npz"))
    meta = json.loads(str(d["meta"]))
    arrays = [d["site%d" % i] for i in range(len(meta["labels"]))]
    return arrays, meta


@pytest.fixture
def fixture_mps_arrays():
    return load_fixture_arrays()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
