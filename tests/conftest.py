import json
import math
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ENTR is a sum of p ln p in fp64 in both implementations, in different
# orders and with different log implementations; stored in the input type.
ENTR_RTOL = {np.float32: 1e-6, np.float64: 1e-12}


def assert_measures(got, exp, dt, what):
    """got/exp: [..., 5] arrays of one measure tuple layout.  DET/LAM, L/TT,
    Lmax/TTmax and RR bit for bit; ENTR within ENTR_RTOL relative."""
    got = np.asarray(got).reshape(-1, 5)
    exp = np.asarray(exp).reshape(-1, 5)
    for k in (0, 1, 2, 4):
        assert same_bits(got[:, k].copy(), exp[:, k].copy()), (what, k, got[:, k], exp[:, k])
    g, e = got[:, 3].astype(np.float64), exp[:, 3].astype(np.float64)
    tol = ENTR_RTOL[dt]
    for a, b in zip(g, e):
        if math.isnan(b):
            assert math.isnan(a), (what, a, b)
        else:
            assert abs(a - b) <= tol * max(abs(b), 1e-300), (what, a, b)


def check_case(amd, oracle, x, taus, embs, lmins, thrs, dist, what):
    """RR, DET and LAM of the GPU against the oracle's."""
    dt = x.dtype.type
    rr = amd.RR(x, taus, embs, thrs, dist)
    rr_o, e = oracle.RR(x, taus, embs, thrs, dist)
    assert e == 0
    assert same_bits(rr.reshape(-1), rr_o), (what, "RR", rr.reshape(-1), rr_o)
    det = amd.DET(x, taus, embs, lmins, thrs, dist)
    det_o, e = oracle.DET(x, taus, embs, lmins, thrs, dist)
    assert e == 0
    assert_measures(det, det_o, dt, what + " DET")
    lam = amd.LAM(x, taus, embs, lmins, thrs, dist)
    lam_o, e = oracle.LAM(x, taus, embs, lmins, thrs, dist)
    assert e == 0
    assert_measures(lam, lam_o, dt, what + " LAM")
