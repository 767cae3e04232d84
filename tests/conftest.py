import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    import numpy as np
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    ins = {k[3:]: d[k] for k in d.files if k.startswith("in_")}
    outs = {k[4:]: d[k] for k in d.files if k.startswith("out_")}
    return ins, outs


@pytest.fixture
def load_golden():
    return golden
