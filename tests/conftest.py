import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "golden_reference.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_rp():
    with open(os.path.join(GOLDEN, "golden_rp.json")) as f:
        return json.load(f)


def rp_case(case):
    """(x, tau, emb, thr, dist, output dtype, expected N_c x N_c plot) of a
    golden accrqa_RP case."""
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    x = unhex(case["x"], dt)
    thr = unhex([case["thr"]], dt)[0]
    ot = np.int32 if case["out"] == "i32" else np.int8
    nc = case["nc"]
    bits = np.unpackbits(np.frombuffer(bytes.fromhex(case["bits"]), dtype=np.uint8))[: nc * nc]
    return x, case["tau"], case["emb"], thr, case["dist"], ot, bits.reshape(nc, nc).astype(ot)


@pytest.fixture(scope="session")
def crqa_kat():
    with open(os.path.join(GOLDEN, "crqa_kat.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def crqa_series():
    with open(os.path.join(GOLDEN, "crqa_input.txt")) as f:
        return np.array([float(v) for v in f.read().split()], dtype=np.float64)


def unhex(lst, dt):
    return np.array([float.fromhex(v) for v in lst], dtype=np.float64).astype(dt)


def dense(sp, nc):
    h = np.zeros(nc + 1, dtype=np.uint64)
    for k, v in sp.items():
        h[int(k)] = v
    return h


def case_arrays(case):
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    return dt, unhex(case["x"], dt), unhex(case["thr"], dt)


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))
