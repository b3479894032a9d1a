"""GPU parity on the benchmark configurations' own generators, sizes and
threshold grids (the oracle wherever it finishes in seconds, closed forms and
size-independent properties beyond that).

* the logistic-map configuration at its full N = 10,000 against the oracle;
* the batched configuration: a sample of its series (all three generators,
  lengths 540 .. 8,364) through the batched call on its full grid
  (tau {1, 2, 4} x m {2, 3, 5} x 8 thresholds x both norms) against the
  oracle, series by series; and all 4,096 series in one batched call per
  norm against the per-series public API on every 16th series;
* the AR(1)+sine configuration at m = 32 and m = 64, N = 262,144: the MFMA
  form, the exact form and the direct (no tile store) form agree bit for bit,
  every histogram satisfies sum l*h_V = sum l*h_D = the separate recurrence
  count, and a prefix matches the oracle's histograms;
* plateau series (tests/plateau.py) of N = 131,072 and 262,144 whose h_D and
  h_V follow in closed form, with lines of every length tier the line scan
  keeps apart (LDS < 128, workgroup-private < 2,048, device atomics above,
  and joined segment carries below and above 4,096), in both norms, both
  distance forms and both split modes.
"""
import numpy as np
import pytest

from accrqa_amd import workloads as W
from conftest import assert_measures, check_case, same_bits, say
from plateau import plateau_hists, plateau_lengths, plateau_series

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import accrqa_amd
    accrqa_amd.load()
    yield accrqa_amd
    accrqa_amd.set_form(accrqa_amd.FORM_AUTO)
    accrqa_amd.set_split(accrqa_amd.SPLIT_AUTO)


def test_config1_full_size_vs_oracle(amd, oracle):
    c = W.config1()
    say("config1 N=10000")
    check_case(amd, oracle, c["x"], c["tau"], c["emb"], c["lmin"] + [1, 3], c["thr"], 1, "config1")
    h = amd.histograms(c["x"], 1, 3, np.asarray(c["thr"], np.float32), 1, which=7)
    for k, t in enumerate(c["thr"]):
        assert np.array_equal(h["lam"][k], oracle.hist("lam", c["x"], np.float32(t), 1, 3, 1)), t
        assert np.array_equal(h["det"][k], oracle.hist("det", c["x"], np.float32(t), 1, 3, 1)), t


# series of the batched configuration: (index, length) 12 (883, AR), 64
# (1,335, logistic), 20 (2,999, noise), 39 (3,790, AR), 145 (5,453, logistic),
# 188 (8,364, noise), 108 (560, AR), 49 (540, logistic)
CONFIG4_SAMPLE = [12, 64, 20, 39, 145, 188, 108, 49]


def test_config4_sample_vs_oracle(amd, oracle):
    c = W.config4(sample=CONFIG4_SAMPLE)
    xs = c["series"]
    thr = np.array([[r * s.std() for r in c["thr_rel"]] for s in xs], dtype=np.float32)
    for d in c["norms"]:
        say("config4 sample, norm", d)
        got = amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, d)
        for k, s in enumerate(xs):
            rr, e = oracle.RR(s, c["tau"], c["emb"], thr[k], d)
            assert e == 0
            assert same_bits(got["RR"][k].reshape(-1), rr), (CONFIG4_SAMPLE[k], d, "RR")
            det, e = oracle.DET(s, c["tau"], c["emb"], c["lmin"], thr[k], d)
            assert e == 0
            assert_measures(got["DET"][k], det, np.float32, f"config4 {CONFIG4_SAMPLE[k]} DET {d}")
            lam, e = oracle.LAM(s, c["tau"], c["emb"], c["lmin"], thr[k], d)
            assert e == 0
            assert_measures(got["LAM"][k], lam, np.float32, f"config4 {CONFIG4_SAMPLE[k]} LAM {d}")


def test_config4_full_batch_vs_single_series(amd):
    """The whole batched configuration (4,096 series, both norms) in one
    batched call per norm, against the per-series public API on every 16th
    series (every (tau, m) block, RR bit for bit, DET/LAM to the parity bar),
    and RR of every series against a second batched call that asks for RR
    only (LAM bands alone)."""
    c = W.config4()
    xs = c["series"]
    thr = np.array([[r * s.std() for r in c["thr_rel"]] for s in xs], dtype=np.float32)
    for d in c["norms"]:
        say("config4 full batch, norm", d)
        got = amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, d)
        assert got["RR"].shape == (len(xs), 3, 3, 8)
        for k in range(0, len(xs), 16):
            rr = amd.RR(xs[k], c["tau"], c["emb"], thr[k], d)
            det = amd.DET(xs[k], c["tau"], c["emb"], c["lmin"], thr[k], d)
            lam = amd.LAM(xs[k], c["tau"], c["emb"], c["lmin"], thr[k], d)
            assert same_bits(got["RR"][k], rr), (k, d)
            assert_measures(got["DET"][k], det, np.float32, f"config4 full {k} DET {d}")
            assert_measures(got["LAM"][k], lam, np.float32, f"config4 full {k} LAM {d}")
        rr_only = amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, d, want=("RR",))
        assert same_bits(rr_only["RR"], got["RR"]), d


def _hists(amd, x, tau, emb, thr, dist, which=7):
    return amd.histograms(x, tau, emb, np.asarray(thr, dtype=x.dtype), dist, which=which)


def _same(a, b, what):
    for key in ("rr", "det", "lam"):
        if key in a and key in b:
            assert np.array_equal(a[key], b[key]), (what, key)


def _counts(h, what):
    nc = h["lam"].shape[1] - 1
    lv = np.arange(nc + 1, dtype=np.uint64)
    for k in range(h["lam"].shape[0]):
        lam = int((lv * h["lam"][k]).sum())
        det = int((lv * h["det"][k]).sum())
        assert lam == int(h["rr"][k]) and det == int(h["rr"][k]), (what, k, lam, det, int(h["rr"][k]))


@pytest.mark.parametrize("m,prefix", [(32, 12000), (64, 8000)])
def test_config3_m32_m64(amd, oracle, m, prefix):
    c = W.config3(m=m)
    x, thr = c["x"], c["thr"]
    amd.set_form(amd.FORM_EXACT)
    say("config3", m, "exact split")
    ex = _hists(amd, x, 1, m, thr, 1)
    amd.set_form(amd.FORM_MFMA)
    say("config3", m, "mfma split")
    mx = _hists(amd, x, 1, m, thr, 1)
    _same(ex, mx, f"config3 m={m} exact vs mfma")
    amd.set_split(amd.SPLIT_NEVER)
    say("config3", m, "mfma direct")
    md = _hists(amd, x, 1, m, thr, 1)
    amd.set_split(amd.SPLIT_AUTO)
    _same(mx, md, f"config3 m={m} split vs direct")
    _counts(mx, f"config3 m={m}")
    # a prefix against the oracle, in the default form
    amd.set_form(amd.FORM_AUTO)
    xp = np.ascontiguousarray(x[:prefix + m - 1])
    say("config3", m, "prefix", prefix)
    hp = _hists(amd, xp, 1, m, thr, 1, which=6)
    for k, t in enumerate(thr):
        assert np.array_equal(hp["lam"][k], oracle.hist("lam", xp, np.float32(t), 1, m, 1)), (m, t)
        assert np.array_equal(hp["det"][k], oracle.hist("det", xp, np.float32(t), 1, m, 1)), (m, t)


# Plateau lengths: runs exactly at the tier edges (127/128, 2,047/2,048,
# 4,095/4,096/4,097 after the embedding's m - 1 is taken off), then
# 2,500 .. 6,000.
def _edge_lengths(m):
    return [B + m - 1 for B in (1, 2, 127, 128, 2047, 2048, 4095, 4096, 4097)]


@pytest.mark.parametrize("n,m,forms", [(131072, 1, ("exact", "mfma")), (262144, 16, ("mfma", "exact"))])
def test_plateau_closed_form_all_tiers(amd, n, m, forms):
    lengths = plateau_lengths(n, 2500, 6000, seed=n + m, extra=_edge_lengths(m))
    x = plateau_series(lengths, np.float32)
    count, hd, hv = plateau_hists(lengths, m)
    assert hv[4096:].sum() > 0 and hv[2048:4096].sum() > 0 and hd[4096:-1].sum() > 0
    thr = np.array([0.0, 0.5, 0.75], np.float32)
    for dist in (1, 2):
        for form in forms:
            if dist == 2 and form == "mfma":
                continue
            amd.set_form(amd.FORM_MFMA if form == "mfma" else amd.FORM_EXACT)
            for split in (amd.SPLIT_AUTO, amd.SPLIT_NEVER):
                amd.set_split(split)
                say("plateau", n, m, dist, form, split)
                h = _hists(amd, x, 1, m, thr, dist)
                for k in range(thr.size):
                    assert int(h["rr"][k]) == count, (n, m, dist, form, split, k)
                    bad_v = np.nonzero(h["lam"][k] != hv)[0]
                    bad_d = np.nonzero(h["det"][k] != hd)[0]
                    assert bad_v.size == 0, (n, m, dist, form, split, k, "h_V bins", bad_v[:8])
                    assert bad_d.size == 0, (n, m, dist, form, split, k, "h_D bins", bad_d[:8])
    amd.set_form(amd.FORM_AUTO)
    amd.set_split(amd.SPLIT_AUTO)
