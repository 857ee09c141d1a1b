import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def load_golden(name):
    import json

    import numpy as np

    g = dict(np.load(os.path.join(GOLDEN_DIR, name + ".npz")))
    g["config"] = json.loads(str(g["config_json"]))
    return g


@pytest.fixture
def golden():
    return load_golden
