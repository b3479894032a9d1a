"""Pins the plateau closed form (tests/plateau.py) on the CPU before the GPU
tests lean on it at sizes the oracle cannot reach: against the oracle's
histograms and recurrence counts for small plateau series (both norms, fp32
and fp64, several embedding dimensions and thresholds in [0, 1)), and the
oracle's own view of what the closed form asserts (the matrix is block
diagonal)."""
import numpy as np
import pytest

from plateau import plateau_hists, plateau_lengths, plateau_series


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dist", [1, 2])
def test_closed_form_matches_oracle(oracle, dt, dist):
    cases = [([5, 9, 3, 12, 7], 1), ([5, 9, 3, 12, 7], 3), ([40, 33, 8, 61, 20, 9], 8),
             (plateau_lengths(700, 17, 90, seed=1, extra=(1, 2, 128, 127)), 1),
             (plateau_lengths(900, 20, 150, seed=2, extra=(129,)), 16)]
    for lengths, m in cases:
        lengths = [L for L in lengths if L >= m]
        x = plateau_series(lengths, dt)
        count, hd, hv = plateau_hists(lengths, m)
        for thr in (0.0, 0.25, 0.5, 0.999):
            t = dt(thr)
            assert np.array_equal(oracle.hist("det", x, t, 1, m, dist), hd), (lengths, m, thr)
            assert np.array_equal(oracle.hist("lam", x, t, 1, m, dist), hv), (lengths, m, thr)
            assert oracle.rr_count(x, t, 1, m, dist) == count, (lengths, m, thr)
        # at a threshold of 1 neighbouring levels recur and the closed form no longer holds
        assert not np.array_equal(oracle.hist("lam", x, dt(1.0), 1, m, dist), hv)


def test_closed_form_small_brute_force():
    """The closed form against a direct recount of a tiny matrix."""
    lengths, m = [4, 6, 3, 5], 2
    x = plateau_series(lengths, np.float64)
    nc = x.size - (m - 1)
    v = np.stack([x[i:i + nc] for i in range(m)], axis=1)
    R = (np.abs(v[:, None, :] - v[None, :, :]).max(axis=2) <= 0.5).astype(int)
    hv = np.zeros(nc + 1, np.uint64)
    hd = np.zeros(nc + 1, np.uint64)

    def runs(seq, h):
        L = 0
        for b in list(seq) + [0]:
            if b:
                L += 1
            elif L:
                h[L] += 1
                L = 0
    for j in range(nc):
        runs(R[:, j], hv)
    for r in range(1, nc):
        runs(np.diagonal(R, r), hd)
    hd *= 2
    hd[nc] += 1
    count, chd, chv = plateau_hists(lengths, m)
    assert count == int(R.sum())
    assert np.array_equal(chd, hd) and np.array_equal(chv, hv)


def test_lengths_helper():
    L = plateau_lengths(131072, 2500, 6000, seed=3, extra=(2048, 4096))
    assert sum(L) == 131072 and L[:2] == [2048, 4096] and min(L) >= 2048
