"""Shared fixtures: the library binding, the oracle and the golden fixtures."""
import importlib.util
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "parallel-randomized-cholesky_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")

if os.path.join(ROOT, "oracle") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    if "rchol_amd" in sys.modules:
        return sys.modules["rchol_amd"]
    spec = importlib.util.spec_from_file_location("rchol_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rchol_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_workloads():
    load_package()
    import importlib
    return importlib.import_module("rchol_amd.workloads")


@pytest.fixture(scope="session")
def rc():
    return load_package()


@pytest.fixture(scope="session")
def wl():
    return load_workloads()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        return json.load(f)


def golden_path(name):
    return os.path.join(GOLDEN, name)


def read_mtx_py(path):
    """Minimal Matrix Market coordinate reader for feeding the oracle in tests
    (symmetric files expanded, pattern -> 1.0, 0-based)."""
    import numpy as np
    with open(path) as f:
        header = f.readline().split()
        field, sym = header[3].lower(), header[4].lower()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, _, m = (int(t) for t in line.split())
        rows, cols, vals = [], [], []
        for _ in range(m):
            t = f.readline().split()
            r, c = int(t[0]) - 1, int(t[1]) - 1
            v = 1.0 if field == "pattern" else float(t[2])
            rows.append(r); cols.append(c); vals.append(v)
            if sym == "symmetric" and r != c:
                rows.append(c); cols.append(r); vals.append(v)
    return n, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals)
