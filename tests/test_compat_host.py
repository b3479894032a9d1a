"""The reference-compatible Python module (accrqa_amd.compat) on the CPU, with
its RR / DET / LAM routed to the oracle: the reference's array shapes (the
DET cube (nTaus, nEmbs, nLmins, nThresholds, 5) over the l_min-fastest
buffer, quirk kept), tidy DataFrame columns and rows (accrqa.py:170-187,
302-327, 442-467), and its argument checks."""
import numpy as np
import pytest

import accrqa_amd.compat as accrqa


@pytest.fixture
def routed(monkeypatch, oracle):
    def rr(x, tau, emb, thr, d):
        out, e = oracle.RR(x, tau, emb, thr, d)
        assert e == 0
        return out

    def dl(kind):
        def f(x, tau, emb, lm, thr, d, ent=True):
            out, e = getattr(oracle, kind)(x, tau, emb, lm, thr, d)
            assert e == 0
            return out
        return f

    monkeypatch.setattr(accrqa, "_RR", rr)
    monkeypatch.setattr(accrqa, "_DET", dl("DET"))
    monkeypatch.setattr(accrqa, "_LAM", dl("LAM"))


def test_shapes_and_tidy_frames(routed, oracle):
    rng = np.random.default_rng(83)
    x = rng.standard_normal(300).astype(np.float32)
    eu = accrqa.accrqaDistance("euclidean")
    tau, emb, lm, thr = [1, 2], [2, 3], [2, 3, 4], [0.3, 0.6]
    rr = accrqa.RR(x, tau, emb, thr, eu, tidy_data=False)
    assert rr.shape == (2, 2, 2)
    det = accrqa.DET(x, tau, emb, lm, thr, eu, tidy_data=False)
    flat, _ = oracle.DET(x, tau, emb, lm, thr, 1)
    assert det.shape == (2, 2, 3, 2, 5) and np.array_equal(det.reshape(-1), np.asarray(flat).reshape(-1))
    lam = accrqa.LAM(x, tau, emb, lm, thr, eu, tidy_data=False)
    assert lam.shape == (2, 2, 3, 2, 5)
    t = accrqa.RR(x, tau, emb, thr, eu)
    assert list(t.columns) == ["Delay", "Embedding", "Threshold", "RR"] and len(t) == 8
    assert t["RR"].to_numpy()[3] == rr[0, 1, 1]
    t = accrqa.DET(x, tau, emb, lm, thr, eu)
    assert list(t.columns) == ["Delay", "Embedding", "Lmin", "Threshold", "DET", "L", "Lmax", "ENTR", "RR"]
    assert len(t) == 24 and t["L"].to_numpy()[5] == det[0, 0, 2, 1, 1]
    t = accrqa.LAM(x, tau, emb, lm, thr, eu)
    assert list(t.columns) == ["Delay", "Embedding", "Vmin", "Threshold", "LAM", "TT", "TTmax", "ENTR", "RR"]
    assert t["TTmax"].to_numpy()[7] == lam[0, 1, 0, 1, 2]


def test_checks():
    x = np.zeros(50, np.float32)
    eu = accrqa.accrqaDistance("euclidean")
    with pytest.raises(ValueError):
        accrqa.accrqaDistance("manhattan")
    with pytest.raises(ValueError):
        accrqa.accrqaCompPlatform("tpu")
    with pytest.raises(TypeError):
        accrqa.RR(x, [1], [2], [0.1], 1)
    with pytest.raises(TypeError):
        accrqa.RR(list(x), [1], [2], [0.1], eu)
    with pytest.raises(RuntimeError):
        accrqa.RR(x, [1], [2], [0.1], eu, accrqa.accrqaCompPlatform("nv_gpu"))
    with pytest.raises(TypeError):
        accrqa.DET(x, [1], [2], [2], [0.1], eu, calculate_ENTR=1)
    with pytest.raises(TypeError):
        accrqa.RP(x, 0, 2, 0.1, eu)
    assert accrqa.accrqaCompPlatform("amd_gpu").get_platform_id() == 2048
