"""The reference's Python interface (src/accrqa/accrqa.py, accrqa_distance.py,
accrqa_compplatform.py) on this library: `import accrqa_amd.compat as accrqa`
in place of `import accrqa`.

Same names, arguments, checks and outputs: accrqaDistance, accrqaCompPlatform
(with "amd_gpu" = 2048 added to "cpu" = 1 and "nv_gpu" = 1024), RR, DET, LAM,
RP and RR_target, with tidy pandas output as in the reference
(accrqa.py:170-187, 302-327, 442-467).  Only "amd_gpu" computes here; the
other platform ids raise, as a library without them would.  The DET cube keeps
the reference's shape (nTaus, nEmbs, nLmins, nThresholds, 5) over the C
layer's l_min-fastest buffer, so its tidy rows are labelled exactly as the
reference labels them (accrqa.py:242 and 302-327).
"""
import numpy as np

from . import DET as _DET, LAM as _LAM, RP as _RP, RR as _RR, RR_target as _RR_target


class accrqaDistance:
    """Distance type by name: "euclidean" (1) or "maximal" (2)
    (src/accrqa/accrqa_distance.py)."""
    distance_types = {1: "euclidean", 2: "maximal"}

    def __init__(self, distance_type: str):
        reverse_map = {v: k for k, v in self.distance_types.items()}
        if distance_type not in reverse_map:
            raise ValueError(f"Unknown distance type: {distance_type}")
        self._id = reverse_map[distance_type]

    def get_distance_id(self) -> int:
        return self._id

    def get_distance_name(self) -> str:
        return self.distance_types.get(self._id, "Unknown")


class accrqaCompPlatform:
    """Computational platform by name (src/accrqa/accrqa_compplatform.py), with
    the AMD GPU added."""
    comp_platforms = {1: "cpu", 1024: "nv_gpu", 2048: "amd_gpu"}

    def __init__(self, comp_platform: str):
        reverse_map = {v: k for k, v in self.comp_platforms.items()}
        if comp_platform not in reverse_map:
            raise ValueError(f"Unknown computational platform: {comp_platform}")
        self._id = reverse_map[comp_platform]

    def get_platform_id(self) -> int:
        return self._id

    def get_platform_name(self) -> str:
        return self.comp_platforms.get(self._id, "Unknown")


def _common(input_data, distance_type, comp_platform, tidy_data):
    if not type(distance_type) is accrqaDistance:
        raise TypeError("Distance must be an instance of accrqaDistance")
    if not type(comp_platform) is accrqaCompPlatform:
        raise TypeError("Computational platform must be an instance of accrqaCompPlatform")
    if comp_platform.get_platform_id() != 2048:
        raise RuntimeError(f"platform {comp_platform.get_platform_name()} is not in this library; use amd_gpu")
    if type(input_data) != np.ndarray:
        raise TypeError("Unknown array type of the input_data")
    if not type(tidy_data) == bool:
        raise TypeError("tidy_data must be bool (False or True)")
    return distance_type.get_distance_id()


def _arr(v, dt):
    return np.asarray(v, dtype=dt).reshape(-1)


def RR(input_data, tau_values, emb_values, threshold_values, distance_type,
       comp_platform=accrqaCompPlatform("amd_gpu"), tidy_data=True):
    """accrqa.RR (accrqa.py:85-187): [nTaus, nEmbs, nThresholds] or a tidy
    DataFrame (Delay, Embedding, Threshold, RR)."""
    d = _common(input_data, distance_type, comp_platform, tidy_data)
    tau, emb = _arr(tau_values, np.intc), _arr(emb_values, np.intc)
    thr = _arr(threshold_values, input_data.dtype)
    if tau.size <= 0 or emb.size <= 0 or thr.size <= 0:
        raise TypeError("Number of delays, embedding or thresholds must be greater than zero.")
    m = _RR(input_data, tau, emb, thr, d).reshape(tau.size, emb.size, thr.size)
    if not tidy_data:
        return m
    import pandas as pd
    return pd.DataFrame([{"Delay": t, "Embedding": e, "Threshold": h, "RR": m[a, b, c]}
                         for a, t in enumerate(tau) for b, e in enumerate(emb) for c, h in enumerate(thr)])


def _dl(kind, input_data, tau_values, emb_values, l_values, threshold_values, distance_type, calculate_ENTR,
        comp_platform, tidy_data):
    d = _common(input_data, distance_type, comp_platform, tidy_data)
    if not type(calculate_ENTR) == bool:
        raise TypeError("calculate_ENTR must be bool (False or True)")
    tau, emb, lm = _arr(tau_values, np.intc), _arr(emb_values, np.intc), _arr(l_values, np.intc)
    thr = _arr(threshold_values, input_data.dtype)
    if tau.size <= 0 or emb.size <= 0 or lm.size <= 0 or thr.size <= 0:
        raise TypeError("Number of delays, embedding, minimal lengths or thresholds must be greater than zero.")
    fn = _DET if kind == "DET" else _LAM
    flat = fn(input_data, tau, emb, lm, thr, d, calculate_ENTR).reshape(-1)
    m = flat.reshape(tau.size, emb.size, lm.size, thr.size, 5)
    if not tidy_data:
        return m
    import pandas as pd
    ln, names = ("Lmin", ("DET", "L", "Lmax", "ENTR", "RR")) if kind == "DET" else \
        ("Vmin", ("LAM", "TT", "TTmax", "ENTR", "RR"))
    rows = []
    for a, t in enumerate(tau):
        for b, e in enumerate(emb):
            for c, l in enumerate(lm):
                for k, h in enumerate(thr):
                    r = {"Delay": t, "Embedding": e, ln: l, "Threshold": h}
                    r.update({nm: m[a, b, c, k, q] for q, nm in enumerate(names)})
                    rows.append(r)
    return pd.DataFrame(rows)


def DET(input_data, tau_values, emb_values, lmin_values, threshold_values, distance_type, calculate_ENTR=True,
        comp_platform=accrqaCompPlatform("amd_gpu"), tidy_data=True):
    """accrqa.DET (accrqa.py:189-327)."""
    return _dl("DET", input_data, tau_values, emb_values, lmin_values, threshold_values, distance_type,
               calculate_ENTR, comp_platform, tidy_data)


def LAM(input_data, tau_values, emb_values, vmin_values, threshold_values, distance_type, calculate_ENTR=True,
        comp_platform=accrqaCompPlatform("amd_gpu"), tidy_data=True):
    """accrqa.LAM (accrqa.py:329-467)."""
    return _dl("LAM", input_data, tau_values, emb_values, vmin_values, threshold_values, distance_type,
               calculate_ENTR, comp_platform, tidy_data)


def RP(input_data, tau, emb, threshold, distance_type):
    """accrqa.RP (accrqa.py:22-82): the int8 N_c x N_c plot."""
    if tau <= 0 or emb <= 0:
        raise TypeError("Delay and embedding must be greater than zero.")
    if type(input_data) != np.ndarray:
        raise TypeError("Unknown array type of the input_data")
    if not type(distance_type) is accrqaDistance:
        raise TypeError("Distance must be an instance of accrqaDistance")
    return _RP(input_data, tau, emb, threshold, distance_type.get_distance_id())


def RR_target(input_data, tau, emb, target_RR, distance_type, epsilon=0.01,
              comp_platform=accrqaCompPlatform("amd_gpu"), max_iter=20, threshold_min=0, threshold_max=10):
    """accrqa.RR_target (accrqa.py:470-594): (threshold, RR)."""
    if not type(distance_type) is accrqaDistance:
        raise TypeError("distance must be an instance of accrqaDistance")
    if not type(comp_platform) is accrqaCompPlatform:
        raise TypeError("Computational platform must be an instance of accrqaCompPlatform")
    if comp_platform.get_platform_id() != 2048:
        raise RuntimeError(f"platform {comp_platform.get_platform_name()} is not in this library; use amd_gpu")
    return _RR_target(input_data, tau, emb, target_RR, distance_type.get_distance_id(), epsilon, max_iter,
                      threshold_min, threshold_max)
