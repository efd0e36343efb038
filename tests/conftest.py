import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")

if REPO not in sys.path:
    sys.path.insert(0, REPO)


def _load_pkg():
    name = "gdlog_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(REPO, "gdlog-multinode_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


gdl = _load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")


def read_edges(name):
    """Tab/space separated pairs of a fixture file -> packed u64 keys in file order."""
    rows = np.loadtxt(os.path.join(DATA, name), dtype=np.int64, ndmin=2)
    return gdl.pack(rows[:, 0], rows[:, 1])


@pytest.fixture(scope="session")
def ctx():
    c = gdl.Context(device=0)
    yield c
    c.close()
