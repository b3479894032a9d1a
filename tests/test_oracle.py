"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* golden fixtures written by the compiled reference CPU path (raw histograms,
  RR / DET / LAM outputs, bit-exact);
* the CRQA known-answer values held by the reference's own tests;
* closed forms (constant series) from the reference's R tests;
* exact-integer brute force: integer-valued series make every squared distance
  an exact small integer, so R(i,j) = [s <= k] for eps = sqrt(k) is known
  without any floating-point reasoning, and run histograms are recounted in
  plain Python;
* invariants (sum l*h_V = sum l*h_D = recurrence count);
* the Lmax bisection against a brute-force "largest l with scan[l] > last";
* ENTR against closed forms.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import case_arrays, dense, same_bits, unhex

# ----------------------------------------------------------------- golden


def test_golden_histograms(oracle, golden):
    n_checked = 0
    for case in golden["cases"]:
        dt, x, _ = case_arrays(case)
        for h in case["hist"]:
            thr = dt(float.fromhex(h["thr"]))
            hd = oracle.hist("det", x, thr, h["tau"], h["emb"], case["dist"])
            hv = oracle.hist("lam", x, thr, h["tau"], h["emb"], case["dist"])
            assert np.array_equal(hd, dense(h["hD"], h["nc"])), (case["name"], h["tau"], h["emb"], h["thr"])
            assert np.array_equal(hv, dense(h["hV"], h["nc"])), (case["name"], h["tau"], h["emb"], h["thr"])
            n_checked += 1
    assert n_checked > 100


def test_golden_api_outputs(oracle, golden):
    for case in golden["cases"]:
        dt, x, thr = case_arrays(case)
        rr, e = oracle.RR(x, case["tau"], case["emb"], thr, case["dist"])
        assert e == 0 and same_bits(rr, unhex(case["RR"], dt)), case["name"]
        det, e = oracle.DET(x, case["tau"], case["emb"], case["lmin"], thr, case["dist"])
        assert e == 0 and same_bits(det, unhex(case["DET"], dt)), case["name"]
        lam, e = oracle.LAM(x, case["tau"], case["emb"], case["lmin"], thr, case["dist"])
        assert e == 0 and same_bits(lam, unhex(case["LAM"], dt)), case["name"]


# ----------------------------------------------------------------- CRQA KATs


@pytest.mark.parametrize("kind", ["DET", "LAM", "RR"])
def test_crqa_known_answers(oracle, crqa_kat, crqa_series, kind):
    for rec in crqa_kat[kind]:
        dist = 1 if rec["norm"] == "euclidean" else 2
        tau, emb, thr = int(rec["tau"]), int(rec["emb"]), rec["thr"]
        if kind == "RR":
            out, e = oracle.RR(crqa_series, [tau], [emb], [thr], dist)
            assert e == 0 and abs(out[0] - rec["CRQA_RR"]) <= 1e-4
            continue
        lm = int(rec["lmin"] if kind == "DET" else rec["vmin"])
        out, e = getattr(oracle, kind)(crqa_series, [tau], [emb], [lm], [thr], dist)
        assert e == 0
        if kind == "DET":
            exp = [rec["CRQA_DET"], rec["CRQA_L"], None, rec["CRQA_ENTR"], rec["CRQA_RR"]]
        else:
            exp = [rec["CRQA_LAM"], rec["CRQA_TT"], None, None, rec["CRQA_RR"]]
        for k, v in enumerate(exp):
            if v is not None:
                assert abs(out[k] - v) <= 1e-4, (rec, k, out[k], v)


# ----------------------------------------------------------------- closed forms


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("tau", [1, 2])
def test_constant_series_closed_forms(oracle, dt, tau):
    # tests/testthat/test-accRQA_DET.R:84-107, test-accRQA_LAM.R:89-110, test-accRQA_RR.R:76-94
    x = np.ones(10, dtype=dt)
    det, e = oracle.DET(x, [tau], [1], [2, 3, 4], [0.5], 1)
    assert e == 0
    det = det.reshape(3, 5)
    for row, lmin in zip(det, (2, 3, 4)):
        assert row[0] == dt(1 - 2 * sum(range(1, lmin)) / 100)
        assert row[2] == 9 and row[4] == 1
    lam, e = oracle.LAM(x, [tau], [1], [2], [0.5], 1)
    assert e == 0 and lam[0] == 1 and lam[2] == 10 and lam[4] == 1
    rr, e = oracle.RR(x, [tau], [1], [0.5], 1)
    assert e == 0 and rr[0] == 1


# ----------------------------------------------------------------- brute force


def runs_hist(seq_list, nc):
    h = np.zeros(nc + 1, dtype=np.uint64)
    for seq in seq_list:
        n = 0
        for v in list(seq) + [0]:
            if v:
                n += 1
            elif n:
                h[n] += 1
                n = 0
    return h


def exact_int_R(x, tau, emb, k, dist):
    """R for integer-valued x: s is an exact integer, eps = sqrt(k)."""
    xi = x.astype(np.int64)
    nc = x.size - (emb - 1) * tau
    E = np.stack([xi[j * tau: j * tau + nc] for j in range(emb)], axis=1)
    D = E[:, None, :] - E[None, :, :]
    if dist == 1:
        return (D * D).sum(-1) <= k
    return np.abs(D).max(-1) <= math.sqrt(k)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dist", [1, 2])
def test_exact_integer_brute_force(oracle, dt, dist):
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(30, 160))
        x = rng.integers(-3, 4, size=n).astype(dt)
        tau = int(rng.integers(1, 4))
        emb = int(rng.integers(1, 5))
        nc = n - (emb - 1) * tau
        for k in (0, 1, 2, 4, 5, 9):
            eps = dt(math.sqrt(k))
            R = exact_int_R(x, tau, emb, k, dist)
            assert oracle.rr_count(x, eps, tau, emb, dist) == int(R.sum())
            assert all(oracle.r_element(x, i, j, eps, tau, emb, dist) == int(R[i, j])
                       for i in range(0, nc, 7) for j in range(0, nc, 5))
            hv = runs_hist(R, nc)
            assert np.array_equal(oracle.hist("lam", x, eps, tau, emb, dist), hv)
            hd_half = runs_hist([np.diagonal(R, r) for r in range(1, nc)], nc)
            hd = 2 * hd_half
            hd[nc] += 1
            assert np.array_equal(oracle.hist("det", x, eps, tau, emb, dist), hd)


def test_histogram_invariants(oracle):
    rng = np.random.default_rng(11)
    for trial in range(6):
        dt = np.float32 if trial % 2 else np.float64
        x = rng.standard_normal(int(rng.integers(100, 400))).astype(dt)
        tau, emb, dist = 1 + trial % 3, 1 + trial % 4, 1 + trial % 2
        thr = dt(rng.uniform(0.3, 2.0))
        cnt = oracle.rr_count(x, thr, tau, emb, dist)
        lv = np.arange(x.size - (emb - 1) * tau + 1, dtype=np.uint64)
        assert int((lv * oracle.hist("lam", x, thr, tau, emb, dist)).sum()) == cnt
        assert int((lv * oracle.hist("det", x, thr, tau, emb, dist)).sum()) == cnt


# ----------------------------------------------------------------- measures


def test_lmax_bisection_matches_brute_force(oracle):
    rng = np.random.default_rng(3)
    for trial in range(400):
        hs = int(rng.integers(2, 60))
        h = np.zeros(hs, dtype=np.uint64)
        for _ in range(int(rng.integers(0, 6))):
            h[int(rng.integers(1, hs))] += np.uint64(rng.integers(1, 4))
        is_det = trial % 2
        if is_det:
            h[hs - 1] = 1 + 2 * h[hs - 1]
        _, scan = oracle.process(h)
        last = int(scan[-1]) if is_det else 0
        above = [l for l in range(hs) if int(scan[l]) > last]
        assert oracle.tertiary(scan, is_det) == (max(above) if above else 0), (h, is_det)


def test_entropy_closed_forms(oracle):
    h = np.array([0, 0, 3, 3, 0, 0], dtype=np.uint64)
    _, scan = oracle.process(h)
    assert oracle.quaternary(h, scan, 2) == pytest.approx(math.log(2), rel=1e-15)
    h = np.array([0, 5, 0, 0], dtype=np.uint64)
    _, scan = oracle.process(h)
    assert oracle.quaternary(h, scan, 1) == 0.0
    assert math.copysign(1, oracle.quaternary(h, scan, 2)) == -1.0  # empty: -0.0 as the reference


# ----------------------------------------------------------------- errors


def test_argument_errors(oracle):
    x = np.ones(10, dtype=np.float32)
    assert oracle.RR(x, [], [1], [0.5], 1)[1] == 2
    assert oracle.DET(x, [1], [1], [2], [], 1)[1] == 2
    assert oracle.LAM(x, [4], [4], [2], [0.5], 1)[1] == 2      # N_c < 1
    assert oracle.DET(x, [1], [2], [10], [0.5], 1)[1] == 2     # l_min > N_c
    f = oracle.L.oracle_RR_f32
    tau = np.array([1], dtype=np.int32)
    out = np.zeros(1, dtype=np.float32)
    thr = np.array([0.5], dtype=np.float32)
    assert f(out.ctypes.data, None, 10, tau.ctypes.data, 1, tau.ctypes.data, 1, thr.ctypes.data, 1, 1) == 4


def test_reference_library_agrees_when_present(oracle):
    from pyoracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref not built here (the reference tree is absent)")
    R = Reference()
    rng = np.random.default_rng(17)
    for trial in range(6):
        dt = np.float32 if trial % 2 else np.float64
        x = rng.standard_normal(300).astype(dt)
        for kind in ("det", "lam"):
            assert np.array_equal(oracle.hist(kind, x, dt(0.8), 2, 3, 1 + trial % 2),
                                  R.hist(kind, x, dt(0.8), 2, 3, 1 + trial % 2))
        a, _ = oracle.DET(x, [1, 2], [2], [2, 3], [0.4, 0.9], 1)
        b, _ = R.DET(x, [1, 2], [2], [2, 3], [0.4, 0.9], 1)
        assert same_bits(a, b)


def test_golden_rp(oracle, golden_rp):
    """accrqa_RP's plots, written by the compiled reference: the restatement
    gives them bit for bit (both norms, char and int output, NaN/inf samples,
    negative / zero / infinite thresholds, ties, tau = 0, m = 16)."""
    from conftest import rp_case
    assert len(golden_rp["cases"]) >= 15
    for case in golden_rp["cases"]:
        x, tau, emb, thr, dist, ot, want = rp_case(case)
        got, e = oracle.RP(x, tau, emb, thr, dist, ot)
        assert e == 0 and got.dtype == want.dtype, case["name"]
        assert np.array_equal(got, want), case["name"]
        assert int(got.sum(dtype=np.int64)) == case["sum"], case["name"]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_rp_invariants(oracle, dt):
    """Properties the plot must have whatever the arithmetic: symmetric
    (RN(a - b) = -RN(b - a)), ones on the diagonal of finite rows for a
    threshold >= 0 and zeros on NaN rows in the Euclidean norm (the maximal
    norm skips NaN components, `if (d > mx)`, so its diagonal is all ones),
    all zero for a negative or NaN threshold, total = the
    recurrence count that the RR fixtures pin, and each entry = R(i, j)."""
    rng = np.random.default_rng(23)
    x = rng.standard_normal(180).astype(dt)
    x[50] = np.nan
    for dist in (1, 2):
        for tau, emb, thr in ((1, 3, 0.7), (2, 2, 0.3), (0, 4, 0.1)):
            rp, e = oracle.RP(x, tau, emb, dt(thr), dist)
            assert e == 0
            assert np.array_equal(rp, rp.T)
            nc = rp.shape[0]
            finite = np.array([np.all(np.isfinite(x[i:i + (emb - 1) * tau + 1:max(tau, 1)])) for i in range(nc)])
            assert np.all(np.diag(rp)[finite] == 1)
            assert np.all(np.diag(rp)[~finite] == (0 if dist == 1 else 1))
            assert int(rp.sum(dtype=np.int64)) == oracle.rr_count(x, dt(thr), tau, emb, dist)
            for i, j in rng.integers(0, nc, size=(40, 2)):
                assert rp[i, j] == oracle.r_element(x, int(i), int(j), dt(thr), tau, emb, dist)
        for thr in (-0.5, np.nan):
            rp, e = oracle.RP(x, 1, 3, dt(thr), dist)
            assert e == 0 and not rp.any()
    rp, e = oracle.RP(x, 1, 200, dt(1.0), 1)
    assert e == 2


def test_reference_rp_agrees_when_present(oracle):
    from pyoracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref not built here (the reference tree is absent)")
    R = Reference()
    rng = np.random.default_rng(29)
    for dt in (np.float32, np.float64):
        x = rng.standard_normal(250).astype(dt)
        for dist in (1, 2):
            a, _ = oracle.RP(x, 2, 3, dt(0.9), dist)
            b, _ = R.RP(x, 2, 3, dt(0.9), dist)
            assert np.array_equal(a, b)
    a, _ = oracle.RP(x, 1, 2, 0.5, 1, np.int32)
    b, _ = R.RP(x, 1, 2, 0.5, 1, np.int32)
    assert np.array_equal(a, b)


def test_rr_target_restatement(oracle):
    """The bisection restatement: a returned threshold inside epsilon really
    has that RR; an interval whose bounds give the same RR returns at the
    first midpoint; running out of iterations returns the midpoint of the
    last interval; each step halves the interval toward the target (RR is
    monotone in the threshold)."""
    from pyoracle import rr_target
    rng = np.random.default_rng(67)
    x = rng.standard_normal(400).astype(np.float32)
    t, r, seen = rr_target(oracle, x, 1, 3, 0.2, 1, 0.005)
    out, _ = oracle.RR(x, [1], [3], [np.float32(t)], 1)
    assert out.reshape(-1)[0] == r and abs(r - 0.2) < 0.005
    assert seen[0] == 5.0 and all(abs(b - a) <= 5.0 / 2 ** k for k, (a, b) in enumerate(zip(seen, seen[1:])))
    t, r, seen = rr_target(oracle, x, 1, 3, 0.2, 1, 0.01, 20, 20.0, 30.0)
    assert seen == [25.0] and t == 25.0 and r == 1.0
    t, r, seen = rr_target(oracle, x, 1, 3, 0.2, 1, 1e-9, 3)
    assert seen == [5.0, 2.5, 1.25] and t in (0.625, 1.875)
    for a, b in zip(seen, seen[1:]):
        ra, _ = oracle.RR(x, [1], [3], [np.float32(a)], 1)
        assert (b < a) == (ra.reshape(-1)[0] >= 0.2)
