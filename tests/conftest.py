import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden(name):
    import json
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def solutions():
    return {k: [int(x) for x in v] for k, v in golden("solutions.json")["vectors"].items()}


def padded_eq(a, b):
    n = max(len(a), len(b))
    return list(a) + [0] * (n - len(a)) == list(b) + [0] * (n - len(b))
