"""Piecewise-constant series whose recurrence histograms follow in closed form.

A series of plateaus with levels 0, 1, 2, ... (plateau k holds L_k samples of
value k) embedded with delay 1 and dimension m gives, for any threshold in
[0, 1) and either norm, a block-diagonal recurrence matrix:

* vector i is constant (every coordinate k) for the L_k - m + 1 indices whose
  window lies inside plateau k; these B_k = L_k - m + 1 vectors are mutually
  recurrent and recur with nothing else (every other vector differs from them
  by at least 1 in some coordinate);
* the m - 1 vectors whose windows straddle a boundary are all different from
  each other and from every constant vector, so each recurs only with itself.

Hence, with N_c = N - m + 1:

* vertical lines (h_V): every column of block k is one run of B_k ones, and a
  straddling vector's column is one run of length 1:
  h_V[B_k] += B_k, h_V[1] += (K - 1)(m - 1);
* diagonal lines (h_D, as exported: upper diagonals doubled plus the line of
  identity): upper diagonal r < B_k crosses block k in one run of B_k - r
  ones, bounded by zeros, so each block adds 2 to h_D[1 .. B_k - 1];
  h_D[N_c] = 1;
* the recurrence count is sum_k B_k^2 + (K - 1)(m - 1).

Plateau lengths spanning 1..6,000 give runs in every tier of the GPU line
scan's histograms (workgroup LDS, workgroup-private slices, device atomics,
and the stitch of segment carries below and above its LDS cap).
"""
import numpy as np


def plateau_series(lengths, dtype=np.float32):
    return np.concatenate([np.full(int(L), float(k)) for k, L in enumerate(lengths)]).astype(dtype)


def plateau_hists(lengths, m):
    """(count, h_D, h_V) of plateau_series(lengths) at tau = 1, dimension m,
    any threshold in [0, 1).  Every plateau must hold at least m samples."""
    lengths = [int(L) for L in lengths]
    assert all(L >= m for L in lengths)
    n = sum(lengths)
    nc = n - (m - 1)
    K = len(lengths)
    hd = np.zeros(nc + 1, dtype=np.uint64)
    hv = np.zeros(nc + 1, dtype=np.uint64)
    count = 0
    for L in lengths:
        B = L - m + 1
        hv[B] += B
        if B > 1:
            hd[1:B] += 2
        count += B * B
    hv[1] += (K - 1) * (m - 1)
    count += (K - 1) * (m - 1)
    hd[nc] += 1
    return count, hd, hv


def plateau_lengths(total, lo, hi, seed, extra=()):
    """Plateau lengths: the `extra` ones first, then draws from U[lo, hi] until
    the total reaches `total` samples; the last one is trimmed to fit (and
    kept at least `lo` long by merging into its predecessor if needed)."""
    rng = np.random.default_rng(seed)
    out = [int(v) for v in extra]
    while sum(out) < total:
        out.append(int(rng.integers(lo, hi + 1)))
    over = sum(out) - total
    out[-1] -= over
    if out[-1] < lo:
        out[-2] += out[-1]
        out.pop()
    assert sum(out) == total
    return out
