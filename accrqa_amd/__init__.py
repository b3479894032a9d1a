"""accrqa_amd -- RR / DET / LAM recurrence quantification on AMD MI355X (gfx950).

Thin ctypes layer over the C ABI of ``libaccrqa_amd.so`` (include/accrqa_amd.h),
which is the drop-in GPU platform for AccRQA's ``accrqa_RR`` / ``accrqa_DET`` /
``accrqa_LAM`` (src/AccRQA_functions.cpp:525-555, 780-810, 945-975).  The same
shared library also exports the reference's C++ entry points with the new
platform value ``PLT_AMD_GPU = 2048`` (include/AccRQA_amd.hpp).

The functions below mirror the reference's Python API (src/accrqa/accrqa.py:85,
189, 329) minus the pandas "tidy" output: they return numpy arrays shaped like
the reference's output cubes.  There is no CPU fallback: if the library is not
built or no GPU is present, every call raises.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["RR", "DET", "LAM", "RP", "rp_device", "RR_target", "rqa_batch", "histograms", "rqa_device", "set_form", "set_split", "AccrqaError", "library_path",
           "load", "PLT_AMD_GPU", "FORM_AUTO", "FORM_EXACT", "FORM_MFMA", "SPLIT_AUTO", "SPLIT_NEVER",
           "DST_EUCLIDEAN", "DST_MAXIMAL"]

PLT_AMD_GPU = 2048
DST_EUCLIDEAN = 1
DST_MAXIMAL = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libaccrqa_amd.so")
_lib = None

ERRORS = {
    0: "success",
    1: "generic runtime error",
    2: "wrong arguments",
    3: "unsupported data type",
    4: "array not allocated",
    5: "memory copy failure host<->device",
    6: "wrong memory location",
    7: "GPU runtime not found",
    8: "could not locate GPU device",
    9: "not enough memory on the device (GPU)",
    10: "other GPU error",
    11: "invalid metric selected",
}


class AccrqaError(RuntimeError):
    def __init__(self, code, detail=""):
        self.code = code
        super().__init__(f"accrqa error {code}: {ERRORS.get(code, 'unknown')}" + (f" ({detail})" if detail else ""))


def library_path():
    return _LIB_PATH


def load():
    """Load libaccrqa_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise AccrqaError(7, f"{_LIB_PATH} is not built; run accrqa_amd.build.build()")
    # PyTorch wheels ship their own HIP runtime under the same soname
    # (libamdhip64.so.7).  Loading torch first makes the dynamic loader bind
    # this library to that one copy; loading ours first would bring in
    # /opt/rocm's runtime and torch would then fail to initialise a second one
    # in the same process.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = C.CDLL(_LIB_PATH)
    P, I, SZ = C.c_void_p, C.c_int, C.c_size_t
    for sfx in ("f32", "f64"):
        f = getattr(L, f"accrqa_amd_RR_{sfx}")
        f.argtypes = [P, P, SZ, P, I, P, I, P, I, I, P]
        f.restype = None
        for kind in ("DET", "LAM"):
            f = getattr(L, f"accrqa_amd_{kind}_{sfx}")
            f.argtypes = [P, P, SZ, P, I, P, I, P, I, P, I, I, I, P]
            f.restype = None
        f = getattr(L, f"accrqa_amd_histograms_{sfx}")
        f.argtypes = [P, SZ, I, I, P, I, I, I, P, P, P, P]
        f.restype = None
        f = getattr(L, f"accrqa_amd_rqa_device_{sfx}")
        f.argtypes = [P, SZ, I, I, P, I, P, I, I, P, P, P, P]
        f.restype = I
        fl = C.c_float if sfx == "f32" else C.c_double
        f = getattr(L, f"accrqa_amd_RP_{sfx}")
        f.argtypes = [P, P, SZ, I, I, fl, I, P]
        f.restype = None
        f = getattr(L, f"accrqa_amd_rp_device_{sfx}")
        f.argtypes = [P, SZ, I, I, fl, I, P, P]
        f.restype = I
    for sfx in ("f32", "f64"):
        f = getattr(L, f"accrqa_amd_rqa_batch_{sfx}")
        f.argtypes = [P, P, I, P, I, P, I, P, I, P, I, I, P, P, P, P]
        f.restype = None
    L.accrqa_amd_rqa_device_static_f32.argtypes = [P, SZ, I, I, P, I, P, I, I, P, P, P, P]
    L.accrqa_amd_rqa_device_static_f32.restype = I
    L.accrqa_amd_batch_options.argtypes = [I, I]
    L.accrqa_amd_batch_options.restype = I
    L.accrqa_amd_batch_report.restype = C.c_char_p
    L.accrqa_amd_fp64_mismatch_f32.argtypes = [P, SZ, I, I, P, I, I, P, P]
    L.accrqa_amd_fp64_mismatch_f32.restype = I
    L.accrqa_amd_profile.argtypes = [I]
    L.accrqa_amd_profile.restype = None
    L.accrqa_amd_profile_read.argtypes = [C.c_char_p, P, P]
    L.accrqa_amd_profile_read.restype = I
    L.accrqa_amd_RP_f64_i32.argtypes = [P, P, SZ, I, I, C.c_double, I, P]
    L.accrqa_amd_RP_f64_i32.restype = None
    L.accrqa_amd_set_form.argtypes = [I]
    L.accrqa_amd_set_form.restype = I
    L.accrqa_amd_set_split.argtypes = [I]
    L.accrqa_amd_set_split.restype = I
    L.accrqa_amd_set_slab_rows.argtypes = [I]
    L.accrqa_amd_set_slab_rows.restype = I
    L.accrqa_amd_mx_stats.argtypes = [I]
    L.accrqa_amd_mx_stats.restype = I
    L.accrqa_amd_mx_stats_read.argtypes = [C.c_void_p]
    L.accrqa_amd_mx_stats_read.restype = I
    L.accrqa_amd_mfma16_probe.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.accrqa_amd_mfma16_probe.restype = I
    L.accrqa_amd_set_rp_wide.argtypes = [I]
    L.accrqa_amd_set_rp_wide.restype = I
    L.accrqa_amd_set_rp_sym.argtypes = [I]
    L.accrqa_amd_set_rp_sym.restype = I
    L.accrqa_amd_set_scan_waves.argtypes = [I]
    L.accrqa_amd_set_scan_waves.restype = I
    L.accrqa_amd_set_overlap.argtypes = [I]
    L.accrqa_amd_set_overlap.restype = I
    L.accrqa_amd_set_gram.argtypes = [I]
    L.accrqa_amd_set_gram.restype = I
    L.accrqa_amd_gram_probe.argtypes = [I, P, P, P, P, P, P]
    L.accrqa_amd_gram_probe.restype = I
    L.accrqa_amd_last_error.restype = C.c_char_p
    L.accrqa_amd_version.restype = C.c_char_p
    _lib = L
    return L


FORM_AUTO, FORM_EXACT, FORM_MFMA = 0, 1, 2


def set_form(form):
    """Euclidean distance form: FORM_AUTO (MFMA Gram + exact recheck for
    m >= 8), FORM_EXACT (exact VALU sums only) or FORM_MFMA (MFMA form wherever
    it applies: fp32 input, and fp64 input in the LAM pass with the Gram of
    the series rounded to fp32 and an fp64 recheck).  The Gram runs on FP16
    matrix cores with each sample split into two halves (hi/lo).  Results are identical; only
    speed differs."""
    e = load().accrqa_amd_set_form(int(form))
    if e:
        raise AccrqaError(e, f"form {form}")


SPLIT_AUTO, SPLIT_NEVER = 0, 1


def set_split(mode):
    """LAM/DET work split: SPLIT_AUTO (rank the upper triangle once, keep the
    tiles and read the rest of LAM and all of DET from them, when they fit in
    device memory) or SPLIT_NEVER (LAM ranks full columns, DET its own pairs).
    Results are identical; only speed and memory differ."""
    e = load().accrqa_amd_set_split(int(mode))
    if e:
        raise AccrqaError(e, f"split mode {mode}")


def set_slab_rows(rows):
    """Slab form of the split: 0 = automatic (slabs only when the tile store
    of all rows does not fit), rows > 0 = always, slabs of that many rows
    (a multiple of 64, at least 256)."""
    e = load().accrqa_amd_set_slab_rows(int(rows))
    if e:
        raise AccrqaError(e, "set_slab_rows")


def mx_stats(enable=True):
    """Switch the MFMA-form decision counters on (zeroing them) or off."""
    e = load().accrqa_amd_mx_stats(int(bool(enable)))
    if e:
        raise AccrqaError(e, "mx_stats")


def mx_stats_read():
    """(elements ranked, flagged by the table, re-summed exactly) since mx_stats(True)."""
    c = (C.c_ulonglong * 3)()
    e = load().accrqa_amd_mx_stats_read(c)
    if e:
        raise AccrqaError(e, "mx_stats_read")
    return int(c[0]), int(c[1]), int(c[2])


def mfma16_probe(a, b, c):
    """c + a @ b by one matrix-core instruction of the MFMA form's Gram
    (diagnostic): a float16 [32, 16], b float16 [16, 32], c float32 [32, 32]."""
    a = np.ascontiguousarray(a, dtype=np.float16)
    b = np.ascontiguousarray(b, dtype=np.float16)
    c = np.ascontiguousarray(c, dtype=np.float32)
    if a.shape != (32, 16) or b.shape != (16, 32) or c.shape != (32, 32):
        raise ValueError("mfma16_probe: shapes (32,16), (16,32), (32,32)")
    d = np.empty((32, 32), dtype=np.float32)
    e = load().accrqa_amd_mfma16_probe(a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data)
    if e:
        raise AccrqaError(e, "mfma16_probe")
    return d


def set_rp_wide(on):
    """Recurrence plot, exact form: four adjacent columns per lane where the
    stores align (default on), or one."""
    e = load().accrqa_amd_set_rp_wide(1 if on else 0)
    if e:
        raise AccrqaError(e, "set_rp_wide")


def set_rp_sym(on):
    """Recurrence plot of bytes, both forms: the upper triangle only, each
    64 x 64 square stored mirrored as well, where N_c % 16 == 0 and the plot
    is 16-byte aligned (default on), or every element.  The plots are
    identical."""
    e = load().accrqa_amd_set_rp_sym(1 if on else 0)
    if e:
        raise AccrqaError(e, "set_rp_sym")


def set_overlap(on):
    """Split form: the DET tile scan beside the LAM scans on a second stream
    (default on), or after them.  Results are identical."""
    e = load().accrqa_amd_set_overlap(1 if on else 0)
    if e:
        raise AccrqaError(e, "set_overlap")


def set_scan_waves(waves):
    """Waves per SIMD of the three tile scans (default 7; 0 = as many as fit).
    A tuning aid: results are identical for every value."""
    e = load().accrqa_amd_set_scan_waves(int(waves))
    if e:
        raise AccrqaError(e, "set_scan_waves")


GRAM_HILO, GRAM_F16, GRAM_BF16, GRAM_FP8 = 0, 1, 2, 3


def set_gram(gram):
    """Gram operand precision of the split MFMA form's LAM rank pass (the
    measured comparison of precisions): GRAM_HILO (default), GRAM_F16,
    GRAM_BF16 or GRAM_FP8; decisions stay exact in every one."""
    e = load().accrqa_amd_set_gram(int(gram))
    if e:
        raise AccrqaError(e, "set_gram")


def gram_probe(gram, a, b, c):
    """One matrix-core instruction of precision gram on fp32 operands a [32,
    16], b [16, 32] and c [32, 32], rounded as the Gram's windows are: returns
    (d = c + qa @ qb from the instruction, qa, qb the rounded operands)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    c = np.ascontiguousarray(c, dtype=np.float32)
    if a.shape != (32, 16) or b.shape != (16, 32) or c.shape != (32, 32):
        raise ValueError("gram_probe: shapes (32,16), (16,32), (32,32)")
    d = np.empty((32, 32), dtype=np.float32)
    qa = np.empty((32, 16), dtype=np.float32)
    qb = np.empty((16, 32), dtype=np.float32)
    e = load().accrqa_amd_gram_probe(int(gram), _p(a), _p(b), _p(c), _p(d), _p(qa), _p(qb))
    if e:
        raise AccrqaError(e, "gram_probe")
    return d, qa, qb


def rqa_device_static(d_input, n, tau, emb, threshold_values, lmin_values, distance_type, d_rr=0, d_det=0, d_lam=0,
                      stream=0):
    """rqa_device for fp32 series without any device allocation (the
    library's fixed work memory; at most 32,768 samples, 32 thresholds, 16
    l_min values).  Calls that overlap on one GPU race."""
    L = load()
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=np.float32).reshape(-1))
    lm = _ints(lmin_values)
    e = L.accrqa_amd_rqa_device_static_f32(C.c_void_p(d_input), n, int(tau), int(emb), _p(thr), thr.size, _p(lm),
                                           lm.size, _dist(distance_type), C.c_void_p(d_rr or None),
                                           C.c_void_p(d_det or None), C.c_void_p(d_lam or None),
                                           C.c_void_p(stream or None))
    if e != 0:
        raise AccrqaError(e, L.accrqa_amd_last_error().decode())


def fp64_mismatches(d_input, n, tau, emb, threshold_values, distance_type="euclidean", stream=0):
    """Diagnostic for fp32 series in device memory (d_input: device pointer):
    per threshold, the pairs of the full matrix whose fp32 decision (the
    reference's arithmetic) differs from the decision in fp64 on the same
    samples.  Returns a uint64 array in the thresholds' order."""
    L = load()
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=np.float32).reshape(-1))
    out = np.zeros(thr.size, dtype=np.uint64)
    e = L.accrqa_amd_fp64_mismatch_f32(C.c_void_p(d_input), n, int(tau), int(emb), _p(thr), thr.size,
                                       _dist(distance_type), _p(out), C.c_void_p(stream or None))
    if e != 0:
        raise AccrqaError(e, L.accrqa_amd_last_error().decode())
    return out


def _dist(d):
    if isinstance(d, str):
        d = {"euclidean": DST_EUCLIDEAN, "maximal": DST_MAXIMAL}.get(d.lower(), 0)
    return int(d)


def _sfx(dt):
    if dt == np.float64:
        return "f64"
    if dt == np.float32:
        return "f32"
    raise AccrqaError(3, f"input dtype {dt} (float32 or float64 expected)")


def _ints(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1))


def _raise(err):
    if err.value != 0:
        L = load()
        raise AccrqaError(err.value, L.accrqa_amd_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def RR(input_data, tau_values, emb_values, threshold_values, distance_type="euclidean"):
    """Recurrence rate, shape [nTaus, nEmbs, nThresholds] (input dtype)."""
    L = load()
    x = np.ascontiguousarray(input_data)
    sfx = _sfx(x.dtype)
    tau, emb = _ints(tau_values), _ints(emb_values)
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=x.dtype).reshape(-1))
    out = np.zeros((tau.size, emb.size, thr.size), dtype=x.dtype)
    err = C.c_int(-1)
    getattr(L, f"accrqa_amd_RR_{sfx}")(_p(out), _p(x), x.size, _p(tau), tau.size, _p(emb), emb.size, _p(thr),
                                       thr.size, _dist(distance_type), C.byref(err))
    _raise(err)
    return out


def rqa_batch(series, tau_values, emb_values, lmin_values, threshold_values, distance_type="euclidean",
              want=("RR", "DET", "LAM")):
    """RR / DET / LAM of many series in one call.  series: a list of 1-D
    arrays of one dtype; threshold_values: [nseries][T] (or one list of T for
    every series).  Returns {"RR": [nseries, nTaus, nEmbs, T],
    "DET": [nseries, nTaus, nEmbs, T, nL, 5], "LAM": [nseries, nTaus, nEmbs,
    nL, T, 5]} for the measures in `want`: the single-series calls' outputs,
    stacked."""
    L = load()
    xs = [np.ascontiguousarray(s) for s in series]
    dt = xs[0].dtype
    if any(s.dtype != dt for s in xs):
        raise TypeError("rqa_batch: every series must have the same dtype")
    sfx = _sfx(dt)
    cat = np.ascontiguousarray(np.concatenate(xs))
    offs = np.zeros(len(xs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([s.size for s in xs])
    tau, emb, lm = _ints(tau_values), _ints(emb_values), _ints(lmin_values)
    thr = np.asarray(threshold_values, dtype=dt)
    if thr.ndim == 1:
        thr = np.broadcast_to(thr, (len(xs), thr.size))
    thr = np.ascontiguousarray(thr)
    if thr.shape[0] != len(xs):
        raise ValueError("rqa_batch: one row of thresholds per series")
    T, ns = thr.shape[1], len(xs)
    out = {}
    if "RR" in want:
        out["RR"] = np.zeros((ns, tau.size, emb.size, T), dtype=dt)
    if "DET" in want:
        out["DET"] = np.zeros((ns, tau.size, emb.size, T, lm.size, 5), dtype=dt)
    if "LAM" in want:
        out["LAM"] = np.zeros((ns, tau.size, emb.size, lm.size, T, 5), dtype=dt)
    err = C.c_int(-1)
    getattr(L, f"accrqa_amd_rqa_batch_{sfx}")(_p(cat), _p(offs), ns, _p(tau), tau.size, _p(emb), emb.size, _p(lm),
                                              lm.size, _p(thr), T, _dist(distance_type),
                                              _p(out["RR"]) if "RR" in out else None,
                                              _p(out["DET"]) if "DET" in out else None,
                                              _p(out["LAM"]) if "LAM" in out else None, C.byref(err))
    _raise(err)
    return out


def batch_options(autotune=None, graphs=None):
    """rqa_batch settings: autotune (time the persistent kernel's
    instantiations on the first large call and keep the fastest) and graphs
    (capture the batched launches in a hipGraph and replay it for repeated
    calls of the same shape).  None leaves a setting unchanged."""
    a = -1 if autotune is None else int(bool(autotune))
    g = -1 if graphs is None else int(bool(graphs))
    e = load().accrqa_amd_batch_options(a, g)
    if e:
        raise AccrqaError(e, "batch_options")


def batch_report():
    """The batched call's tuning choices and graph builds/replays (text)."""
    return load().accrqa_amd_batch_report().decode()


def RP(input_data, tau, emb, threshold, distance_type="euclidean", out_dtype=np.int8):
    """Recurrence plot, int8 [N_c, N_c] (the reference's char output; int32
    for fp64 input with out_dtype=np.int32), N_c = N - (emb - 1) tau."""
    L = load()
    x = np.ascontiguousarray(input_data)
    sfx = _sfx(x.dtype)
    nc = x.size - (int(emb) - 1) * int(tau)
    i32 = np.dtype(out_dtype) == np.int32
    if i32 and sfx != "f64":
        raise TypeError("int32 output is for fp64 input (as the reference's accrqa_RP overloads)")
    out = np.zeros((max(nc, 0), max(nc, 0)), dtype=np.int32 if i32 else np.int8)
    err = C.c_int(-1)
    fn = L.accrqa_amd_RP_f64_i32 if i32 else getattr(L, f"accrqa_amd_RP_{sfx}")
    fn(_p(out), _p(x), x.size, int(tau), int(emb), float(threshold), _dist(distance_type), C.byref(err))
    _raise(err)
    return out


def rp_device(d_input, n, tau, emb, threshold, distance_type, d_output, stream=0, dtype=np.float32):
    """Device-resident recurrence plot: d_input (n samples) and d_output
    (N_c * N_c bytes) are device pointers; asynchronous on `stream`."""
    L = load()
    sfx = _sfx(np.dtype(dtype))
    e = getattr(L, f"accrqa_amd_rp_device_{sfx}")(C.c_void_p(d_input), n, int(tau), int(emb), float(threshold),
                                                  _dist(distance_type), C.c_void_p(d_output),
                                                  C.c_void_p(stream or None))
    if e != 0:
        raise AccrqaError(e, L.accrqa_amd_last_error().decode())


def _dl(kind, input_data, tau_values, emb_values, l_values, threshold_values, distance_type, calculate_ENTR):
    L = load()
    x = np.ascontiguousarray(input_data)
    sfx = _sfx(x.dtype)
    tau, emb, lm = _ints(tau_values), _ints(emb_values), _ints(l_values)
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=x.dtype).reshape(-1))
    if kind == "DET":
        shape = (tau.size, emb.size, thr.size, lm.size, 5)
    else:
        shape = (tau.size, emb.size, lm.size, thr.size, 5)
    out = np.zeros(shape, dtype=x.dtype)
    err = C.c_int(-1)
    getattr(L, f"accrqa_amd_{kind}_{sfx}")(_p(out), _p(x), x.size, _p(tau), tau.size, _p(emb), emb.size, _p(lm),
                                           lm.size, _p(thr), thr.size, _dist(distance_type), int(bool(calculate_ENTR)),
                                           C.byref(err))
    _raise(err)
    return out


def DET(input_data, tau_values, emb_values, lmin_values, threshold_values, distance_type="euclidean",
        calculate_ENTR=True):
    """DET, L, Lmax, ENTR, RR; shape [nTaus, nEmbs, nThresholds, nLmins, 5] (the reference CPU layout)."""
    return _dl("DET", input_data, tau_values, emb_values, lmin_values, threshold_values, distance_type,
               calculate_ENTR)


def LAM(input_data, tau_values, emb_values, vmin_values, threshold_values, distance_type="euclidean",
        calculate_ENTR=True):
    """LAM, TT, TTmax, ENTR, RR; shape [nTaus, nEmbs, nVmins, nThresholds, 5]."""
    return _dl("LAM", input_data, tau_values, emb_values, vmin_values, threshold_values, distance_type,
               calculate_ENTR)


def histograms(input_data, tau, emb, threshold_values, distance_type="euclidean", which=7):
    """Raw integer results for one (tau, emb): dict with 'rr' [T] (full-matrix
    recurrence counts), 'det' [T, N_c+1] (h_D) and 'lam' [T, N_c+1] (h_V)."""
    L = load()
    x = np.ascontiguousarray(input_data)
    sfx = _sfx(x.dtype)
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=x.dtype).reshape(-1))
    nc = x.size - (int(emb) - 1) * int(tau)
    hs = max(nc + 1, 1)
    rr = np.zeros(thr.size, dtype=np.uint64)
    hd = np.zeros((thr.size, hs), dtype=np.uint64)
    hv = np.zeros((thr.size, hs), dtype=np.uint64)
    err = C.c_int(-1)
    getattr(L, f"accrqa_amd_histograms_{sfx}")(_p(x), x.size, int(tau), int(emb), _p(thr), thr.size,
                                               _dist(distance_type), int(which), _p(rr), _p(hd), _p(hv), C.byref(err))
    _raise(err)
    out = {}
    if which & 1:
        out["rr"] = rr
    if which & 2:
        out["det"] = hd
    if which & 4:
        out["lam"] = hv
    return out


def rqa_device(d_input, n, tau, emb, threshold_values, lmin_values, distance_type, d_rr=0, d_det=0, d_lam=0,
               stream=0, dtype=np.float32):
    """Device-resident RR+DET+LAM for one (tau, emb).  d_* are device pointers
    (ints, e.g. torch tensor .data_ptr()); outputs use the layouts documented in
    include/accrqa_amd.h.  Raises on error."""
    L = load()
    sfx = _sfx(np.dtype(dtype))
    thr = np.ascontiguousarray(np.asarray(threshold_values, dtype=dtype).reshape(-1))
    lm = _ints(lmin_values)
    e = getattr(L, f"accrqa_amd_rqa_device_{sfx}")(C.c_void_p(d_input), n, int(tau), int(emb), _p(thr), thr.size,
                                                   _p(lm), lm.size, _dist(distance_type), C.c_void_p(d_rr or None),
                                                   C.c_void_p(d_det or None), C.c_void_p(d_lam or None),
                                                   C.c_void_p(stream or None))
    if e != 0:
        raise AccrqaError(e, L.accrqa_amd_last_error().decode())


from .rr_target import RR_target  # noqa: E402
