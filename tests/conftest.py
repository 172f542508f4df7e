import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbri17_amd.so on the device)")


@pytest.fixture(scope="session")
def fixtures():
    """The reference's Maxima Ke/Be matrices and test parameters
    (tests/test_bri17.cpp:319-333,344-356,372-531,547-552,570-598)."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_fixtures.json")) as f:
        d = json.load(f)
    for D in (2, 3):
        nd = D * (1 << D)
        d["Ke%d" % D] = np.array(d["Ke%d" % D]).reshape(nd, nd)
        d["Be%d" % D] = np.array(d["Be%d" % D]).reshape(3 if D == 2 else 6, nd)
        d["L%d" % D] = [n * s for n, s in zip(d["shape%d" % D], d["spacing%d" % D])]
    return d


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")
