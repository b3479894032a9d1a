"""RR_target: the threshold whose recurrence rate is within epsilon of a target.

Mirrors the reference's Python function RR_target (src/accrqa/accrqa.py:470-594):
the same arguments, checks, bisection arithmetic (Python floats for the
interval, float32 thresholds for the RR evaluations), stopping rules, warning
and return value (threshold_mid, RR at the last evaluated threshold).  Kept
as in the reference: the three range diagnostics are built but not raised
(accrqa.py:554-559), the interval-width test uses the RRs of the initial
bounds throughout, and a non-float32 series is refused (the reference passes
float32 thresholds, and its binding requires the series and the thresholds
to share a dtype, src/AccRQA_python_bindings.cpp:198-201).

On the GPU the bisection is speculative: the thresholds of the next `depth`
levels of the bisection tree (2^depth - 1 of them) go to the GPU in one
multi-threshold RR pass, which ranks every pair once for all of them, and the
search then walks that tree with the reference's comparisons.  The path, the
thresholds evaluated on it and the result are the sequential loop's; the
number of passes over the N_c^2 pairs drops from max_iter to about
max_iter / depth.
"""
import numpy as np


def _mid(lo, hi):
    return (hi - lo) / 2.0 + lo


def RR_target(input_data, tau, emb, target_RR, distance_type="euclidean", epsilon=0.01, max_iter=20,
              threshold_min=0, threshold_max=10, depth=5):
    from . import RR, _dist
    if type(input_data) != np.ndarray:
        raise TypeError("Unknown array type")
    if tau <= 0 or emb <= 0:
        raise TypeError("Delay and embedding must be greater than zero.")
    if target_RR < 0.0 or target_RR > 1.0:
        raise TypeError("target_RR must be between 0 and 1.")
    if epsilon <= 0.0:
        raise TypeError("epsilon must be positive and non-zero.")
    if max_iter <= 0:
        raise TypeError("max_iter must be positive.")
    if threshold_min >= threshold_max:
        raise TypeError("threshold_min must be smaller then threshold_max.")
    if input_data.dtype != np.float32:
        raise TypeError("RR_target evaluates float32 thresholds, so the series must be float32 "
                        "(the input and threshold values must have the same data type)")
    if depth < 1:
        raise ValueError("depth must be at least 1")
    dist = _dist(distance_type)
    tv, ev = [int(tau)], [int(emb)]

    def rr(thresholds):
        return RR(input_data, tv, ev, np.array(thresholds, dtype=np.float32), dist).flatten()

    first = rr([threshold_min, _mid(threshold_min, threshold_max), threshold_max])
    low_RR, hgh_RR = first[0], first[2]
    lo, hi = threshold_min, threshold_max
    current_RR = None
    it = 0
    while it < max_iter:
        levels = min(depth, max_iter - it)
        # the next `levels` levels of the bisection tree, breadth first: node i
        # of a level has children 2i (RR >= target: max = mid) and 2i + 1 (min = mid)
        tree, nodes = [], [(lo, hi)]
        for _ in range(levels):
            lvl = [(a, b, _mid(a, b)) for a, b in nodes]
            tree.append(lvl)
            nodes = [c for a, b, m in lvl for c in ((a, m), (m, b))]
        vals = rr([m for lvl in tree for _, _, m in lvl])
        off, k = 0, 0
        for lvl in tree:
            it += 1
            threshold_mid = lvl[k][2]
            current_RR = vals[off + k]
            if abs(current_RR - target_RR) < epsilon:
                return threshold_mid, current_RR
            if (hgh_RR - low_RR) < epsilon:
                return threshold_mid, current_RR
            if current_RR < target_RR:
                lo, k = threshold_mid, 2 * k + 1
            else:
                hi, k = threshold_mid, 2 * k
            off += len(lvl)
    print(f"Warning: Bisection search stopped after {max_iter} iterations")
    return _mid(lo, hi), current_RR
