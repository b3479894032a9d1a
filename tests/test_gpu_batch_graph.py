"""The batched call's hipGraph replay and kernel autotuning.

Repeated calls with the same shapes and thresholds replay a captured graph
(k_batch and k_batch_fin per chunk, the histograms zeroed by a kernel): every
replay, with other library work and synchronisations between replays, new
series values in some of them, and calls of another graph (the other norm)
that overwrite the shared descriptor buffer in between, must equal the eager
launches and the single-series calls (RR bit for bit, DET/LAM to the parity
bar).  The first large call times the persistent kernel's instantiations and
reports them.
"""
import numpy as np
import pytest

from accrqa_amd import workloads as W
from conftest import assert_measures, same_bits, say

pytestmark = pytest.mark.gpu

# series of the batched configuration with 1,300 .. 8,400 samples (enough
# pair decisions in the first chunk for the autotuner)
SAMPLE = [188, 200, 221, 145, 169, 259, 39, 285, 288, 20, 26, 32, 64, 73]


def _single(amd, xs, c, thr, d, k):
    rr = amd.RR(xs[k], c["tau"], c["emb"], thr[k], d)
    det = amd.DET(xs[k], c["tau"], c["emb"], c["lmin"], thr[k], d)
    lam = amd.LAM(xs[k], c["tau"], c["emb"], c["lmin"], thr[k], d)
    return rr, det, lam


def _graph_counts(rep):
    """(graphs built, graph replays) from the report's last line."""
    last = rep.strip().split("\n")[-1]
    assert last.startswith("graph:"), rep
    return [int(v) for v in last.replace("graph:", "").replace("built,", "").replace("replays", "").split()]


def _same(a, b):
    return all(same_bits(a[k], b[k]) for k in ("RR", "DET", "LAM"))


def test_batch_graph_replay_and_tuning():
    import accrqa_amd as amd
    amd.load()
    c = W.config4(sample=SAMPLE)
    xs = c["series"]
    thr = np.array([[r * s.std() for r in c["thr_rel"]] for s in xs], dtype=np.float32)
    other = np.cumsum(np.random.default_rng(5).standard_normal(3000)).astype(np.float32)
    try:
        amd.batch_options(autotune=True, graphs=False)
        say("eager reference (and tuning)")
        ref = amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, 1)
        rep = amd.batch_report()
        say(rep)
        assert "k_batch<f32,euclidean,8>" in rep
        for k in (0, 5, 13):
            rr, det, lam = _single(amd, xs, c, thr, 1, k)
            assert same_bits(ref["RR"][k], rr)
            assert_measures(ref["DET"][k], det, np.float32, f"eager DET {k}")
            assert_measures(ref["LAM"][k], lam, np.float32, f"eager LAM {k}")
        amd.batch_options(graphs=True)
        outs = []
        for rep_i in range(4):
            say("graph call", rep_i)
            outs.append(amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, 1))
            amd.DET(other, [2], [3], [2], [0.3, 0.6], 1)   # eager work between replays
        for o in outs:
            assert _same(o, ref)
        rep = amd.batch_report()
        say(rep)
        built, replays = _graph_counts(rep)
        assert built >= 1 and replays >= 3, rep
        # another graph (the maximal norm) overwrites the shared descriptors;
        # both graphs stay kept, and each replay must still see its own
        ref2 = None
        for rep_i in range(3):
            o2 = amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, 2)
            if ref2 is None:
                ref2 = o2
            assert _same(o2, ref2)
            assert _same(amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, 1), ref)
        for k in (2, 9):
            rr, det, lam = _single(amd, xs, c, thr, 2, k)
            assert same_bits(ref2["RR"][k], rr)
            assert_measures(ref2["DET"][k], det, np.float32, f"graph max-norm DET {k}")
            assert_measures(ref2["LAM"][k], lam, np.float32, f"graph max-norm LAM {k}")
        rep = amd.batch_report()
        say(rep)
        built2, replays2 = _graph_counts(rep)
        assert built2 - built <= 1 and replays2 - replays >= 4, rep
        # new series values, same shapes and thresholds: a replay on new data
        xs2 = [np.ascontiguousarray(s[::-1]) for s in xs]
        got = amd.rqa_batch(xs2, c["tau"], c["emb"], c["lmin"], thr, 1)
        for k in (1, 7, 12):
            rr, det, lam = _single(amd, xs2, c, thr, 1, k)
            assert same_bits(got["RR"][k], rr)
            assert_measures(got["DET"][k], det, np.float32, f"replay DET {k}")
            assert_measures(got["LAM"][k], lam, np.float32, f"replay LAM {k}")
        amd.batch_options(graphs=False)
        assert _same(amd.rqa_batch(xs2, c["tau"], c["emb"], c["lmin"], thr, 1), got)
    finally:
        amd.batch_options(autotune=True, graphs=True)


def test_batch_many_thresholds_and_lmins(oracle):
    """The batched kernels with several threshold groups per block (40
    thresholds, some never recurring, duplicates), six l_min values (two
    passes of the finaliser's l_min loop), both norms, fp32 and fp64, series
    from 40 to 5,000 samples, against the single-series calls and the oracle."""
    import accrqa_amd as amd
    amd.load()
    rng = np.random.default_rng(211)
    lens = [40, 300, 1200, 5000]
    taus, embs, lm = [1, 2], [3, 4], [1, 2, 3, 5, 8, 13]
    for dt in (np.float32, np.float64):
        xs = [np.cumsum(rng.standard_normal(n)).astype(dt) for n in lens]
        base = np.concatenate([np.linspace(0.02, 1.5, 34), [-1.0, np.nan, 0.0, 0.3, 0.3, 2.0]])
        thr = np.array([base * s.std() for s in xs], dtype=dt)
        for dist in (1, 2):
            say("batch many", dt.__name__, dist)
            got = amd.rqa_batch(xs, taus, embs, lm, thr, dist)
            for k, s in enumerate(xs):
                rr = amd.RR(s, taus, embs, thr[k], dist)
                det = amd.DET(s, taus, embs, lm, thr[k], dist)
                lam = amd.LAM(s, taus, embs, lm, thr[k], dist)
                assert same_bits(got["RR"][k], rr), (dt, dist, k)
                assert_measures(got["DET"][k], det, dt, f"many DET {dt.__name__} {dist} {k}")
                assert_measures(got["LAM"][k], lam, dt, f"many LAM {dt.__name__} {dist} {k}")
            det_o, e = oracle.DET(xs[2], taus, embs, lm, thr[2], dist)
            assert e == 0
            assert_measures(got["DET"][2], det_o, dt, f"many DET oracle {dt.__name__} {dist}")
            lam_o, e = oracle.LAM(xs[2], taus, embs, lm, thr[2], dist)
            assert e == 0
            assert_measures(got["LAM"][2], lam_o, dt, f"many LAM oracle {dt.__name__} {dist}")
