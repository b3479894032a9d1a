"""The reference's own entry points with the AMD platform branch added.

oracle/_ref/libaccrqa_ref_amd.so is the reference's src/AccRQA_functions.cpp,
AccRQA_python_bindings.cpp and AccRQA_utilities_mem.cpp compiled with the
edits INTEGRATION.md describes (PLT_AMD_GPU = 2048 in the enum, a first
branch in each public overload that calls this library's C ABI, platform
2048 accepted by the Python bindings' check_comp_platform), linked against
accrqa_amd/libaccrqa_amd.so; oracle/ref_amd.py builds it where the reference
tree exists.  In one process the same reference functions run with PLT_CPU
(the reference's own OpenMP path) and with PLT_AMD_GPU (this engine): the
outputs must agree to the parity bar, through the C++ API and through the C
entry points and Mem marshalling the reference's Python package uses.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT, assert_measures, same_bits, say

pytestmark = pytest.mark.gpu

REF_AMD = os.path.join(ROOT, "oracle", "_ref", "libaccrqa_ref_amd.so")
P, I = C.c_void_p, C.c_int

SYM = {
    ("RR", "f32"): "_Z9accrqa_RRPfS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
    ("RR", "f64"): "_Z9accrqa_RRPdS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
    ("DET", "f32"): "_Z10accrqa_DETPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    ("DET", "f64"): "_Z10accrqa_DETPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    ("LAM", "f32"): "_Z10accrqa_LAMPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    ("LAM", "f64"): "_Z10accrqa_LAMPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
}


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_AMD):
        pytest.skip("oracle/_ref/libaccrqa_ref_amd.so is built only where the reference tree exists")
    import accrqa_amd
    accrqa_amd.load()   # torch's HIP runtime and the engine first, then the reference library on top
    L = C.CDLL(REF_AMD)
    for (kind, sfx), name in SYM.items():
        f = getattr(L, name)
        if kind == "RR":
            f.argtypes = [P, P, C.c_size_t, P, I, P, I, P, I, I, I, P]
        else:
            f.argtypes = [P, P, C.c_size_t, P, I, P, I, P, I, P, I, I, I, I, P]
        f.restype = None
    L.mem_create_wrapper.argtypes = [P, I, I, C.c_int32, P, P, P]
    L.mem_create_wrapper.restype = P
    L.mem_free.argtypes = [P]
    L.mem_free.restype = None
    for kind in ("RR", "DET", "LAM"):
        f = getattr(L, f"py_accrqa_{kind}")
        f.argtypes = [P] * 5 + [I, I, P] if kind == "RR" else [P] * 6 + [I, I, I, P]
        f.restype = None
    return L


def _ptr(a):
    return a.ctypes.data


def cxx_call(L, kind, x, taus, embs, lms, thr, dist, platform):
    sfx = "f64" if x.dtype == np.float64 else "f32"
    tau, emb, lm = (np.ascontiguousarray(v, dtype=np.int32) for v in (taus, embs, lms))
    thr = np.ascontiguousarray(thr, dtype=x.dtype)
    n = tau.size * emb.size * thr.size * (1 if kind == "RR" else 5 * lm.size)
    out = np.full(n, -7, dtype=x.dtype)
    err = C.c_int(-1)
    f = getattr(L, SYM[(kind, sfx)])
    if kind == "RR":
        f(_ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size, _ptr(thr), thr.size, dist,
          platform, C.byref(err))
    else:
        f(_ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size, _ptr(lm), lm.size, _ptr(thr),
          thr.size, dist, 1, platform, C.byref(err))
    return out, err.value


MEM_INT, MEM_FLOAT, MEM_DOUBLE, MEM_CPU = 2, 4, 8, 0


def _mem(L, a, keep):
    t = {np.dtype(np.int32): MEM_INT, np.dtype(np.float32): MEM_FLOAT, np.dtype(np.float64): MEM_DOUBLE}[a.dtype]
    shape = np.ascontiguousarray(a.shape, dtype=np.int64)
    keep.append(shape)
    st = C.c_int(0)
    m = L.mem_create_wrapper(_ptr(a), t, MEM_CPU, a.ndim, _ptr(shape), None, C.byref(st))
    assert st.value == 0 and m
    return m


def py_call(L, kind, x, taus, embs, lms, thr, dist, platform):
    """The C entry point the reference's Python package calls
    (src/AccRQA_python_bindings.cpp), with Mem wrappers as it builds them."""
    tau, emb, lm = (np.ascontiguousarray(v, dtype=np.int32) for v in (taus, embs, lms))
    thr = np.ascontiguousarray(thr, dtype=x.dtype)
    if kind == "RR":
        out = np.full((tau.size, emb.size, thr.size), -7, dtype=x.dtype)
    elif kind == "DET":
        out = np.full((tau.size, emb.size, thr.size, lm.size, 5), -7, dtype=x.dtype)
    else:
        out = np.full((tau.size, emb.size, lm.size, thr.size, 5), -7, dtype=x.dtype)
    keep = []
    mems = [_mem(L, a, keep) for a in ((out, x, tau, emb, thr) if kind == "RR" else (out, x, tau, emb, lm, thr))]
    err = C.c_int(0)
    f = getattr(L, f"py_accrqa_{kind}")
    if kind == "RR":
        f(*mems, dist, platform, C.byref(err))
    else:
        f(*mems, dist, 1, platform, C.byref(err))
    for m in mems:
        L.mem_free(m)
    return out.reshape(-1), err.value


def test_reference_dispatch_amd_branch(ref):
    rng = np.random.default_rng(97)
    w = np.cumsum(rng.standard_normal(2600))
    for dt in (np.float32, np.float64):
        x = ((w - w.mean()) / w.std()).astype(dt)
        taus, embs, lms, thr = [1, 3], [2, 9], [2, 4], [0.15, 0.4, 0.9, -1.0]
        for dist in (1, 2):
            for kind in ("RR", "DET", "LAM"):
                say("reference dispatch", dt.__name__, dist, kind)
                cpu, e1 = cxx_call(ref, kind, x, taus, embs, lms, thr, dist, 1)
                amd, e2 = cxx_call(ref, kind, x, taus, embs, lms, thr, dist, 2048)
                assert e1 == 0 and e2 == 0, (kind, e1, e2)
                if kind == "RR":
                    assert same_bits(amd, cpu), (dt, dist, kind)
                else:
                    assert_measures(amd, cpu, dt, f"reference {kind} {dt.__name__} {dist}")
                # the Python package's C entry points, platform 2048
                pyo, e3 = py_call(ref, kind, x, taus, embs, lms, thr, dist, 2048)
                assert e3 == 0, (kind, e3)
                if kind == "RR":
                    assert same_bits(pyo, cpu), (dt, dist, kind, "python binding")
                else:
                    assert_measures(pyo, cpu, dt, f"python binding {kind} {dt.__name__} {dist}")


def test_reference_dispatch_errors(ref):
    """Errors keep the reference's codes on the new platform: no series (2),
    a NULL array (4), N_c < 1 (2, where the reference itself would read out of
    bounds), and an unknown platform in the Python bindings (2)."""
    x = np.ones(10, np.float32)
    out, e = cxx_call(ref, "DET", x, [1], [20], [2], [0.5], 1, 2048)
    assert e == 2
    f = getattr(ref, SYM[("RR", "f32")])
    tau = np.array([1], np.int32)
    err = C.c_int(-1)
    f(None, _ptr(x), x.size, _ptr(tau), 1, _ptr(tau), 1, _ptr(x), 1, 1, 2048, C.byref(err))
    assert err.value == 4
    _, e = py_call(ref, "RR", x, [1], [2], [2], [0.5], 1, 4096)
    assert e == 2
