"""ctypes bindings for the test-only checker libraries under oracle/.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg -- never by the product package (accrqa_amd).

* ``Oracle``    -- liboracle.so, the C restatement of the reference CPU path.
* ``Reference`` -- _ref/libaccrqa_ref.so, the reference's own CPU path compiled
  from its sources (present only where it was built; see oracle/Makefile).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libaccrqa_ref.so")

_P = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t
_I64 = C.c_int64


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _ints(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1))


def _flt(v, dt):
    return np.ascontiguousarray(np.asarray(v, dtype=dt).reshape(-1))


def nc_of(n, tau, emb):
    return int(n) - (int(emb) - 1) * int(tau)


class Oracle:
    """The C restatement (oracle/oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        L = C.CDLL(path)
        self.L = L
        for sfx, fl in (("f32", C.c_float), ("f64", C.c_double)):
            f = getattr(L, f"oracle_R_element_{sfx}")
            f.argtypes = [_P, _I64, _I64, fl, _I, _I, _I]
            f.restype = _I
            f = getattr(L, f"oracle_rr_count_{sfx}")
            f.argtypes = [_P, _SZ, fl, _I, _I, _I]
            f.restype = _I64
            for ot in ("u8", "i32"):
                f = getattr(L, f"oracle_RP_{ot}_{sfx}")
                f.argtypes = [_P, _P, _SZ, fl, _I, _I, _I]
                f.restype = _I
            for kind in ("det", "lam"):
                f = getattr(L, f"oracle_{kind}_hist_{sfx}")
                f.argtypes = [_P, _P, _SZ, fl, _I, _I, _I]
                f.restype = None
            f = getattr(L, f"oracle_RR_{sfx}")
            f.argtypes = [_P, _P, _SZ, _P, _I, _P, _I, _P, _I, _I]
            f.restype = _I
            for kind in ("DET", "LAM"):
                f = getattr(L, f"oracle_{kind}_{sfx}")
                f.argtypes = [_P, _P, _SZ, _P, _I, _P, _I, _P, _I, _P, _I, _I]
                f.restype = _I
        L.oracle_process_hist.argtypes = [_P, _P, _P, _SZ]
        L.oracle_tertiary.argtypes = [_P, _SZ, _I]
        L.oracle_tertiary.restype = C.c_double
        L.oracle_quaternary.argtypes = [_P, _P, _SZ, _SZ]
        L.oracle_quaternary.restype = C.c_double
        L.oracle_num_threads.restype = _I

    @staticmethod
    def _sfx(x):
        return "f64" if np.asarray(x).dtype == np.float64 else "f32"

    def num_threads(self):
        return self.L.oracle_num_threads()

    def r_element(self, x, i, j, thr, tau, emb, dist):
        x = np.ascontiguousarray(x)
        return getattr(self.L, f"oracle_R_element_{self._sfx(x)}")(_ptr(x), i, j, thr, tau, emb, dist)

    def rr_count(self, x, thr, tau, emb, dist):
        x = np.ascontiguousarray(x)
        return getattr(self.L, f"oracle_rr_count_{self._sfx(x)}")(_ptr(x), x.size, thr, tau, emb, dist)

    def RP(self, x, tau, emb, thr, dist, out_dtype=np.int8):
        """(N_c x N_c recurrence plot, error code): int8 (the reference's char
        output) or int32."""
        x = np.ascontiguousarray(x)
        nc = max(nc_of(x.size, tau, emb), 0)
        ot = "i32" if np.dtype(out_dtype) == np.int32 else "u8"
        out = np.zeros((nc, nc), dtype=np.int32 if ot == "i32" else np.int8)
        e = getattr(self.L, f"oracle_RP_{ot}_{self._sfx(x)}")(_ptr(out), _ptr(x), x.size, thr, tau, emb, dist)
        return out, e

    def hist(self, kind, x, thr, tau, emb, dist):
        """Raw histogram h_D ('det') or h_V ('lam'), uint64[N_c+1]."""
        x = np.ascontiguousarray(x)
        h = np.zeros(nc_of(x.size, tau, emb) + 1, dtype=np.uint64)
        getattr(self.L, f"oracle_{kind}_hist_{self._sfx(x)}")(_ptr(h), _ptr(x), x.size, thr, tau, emb, dist)
        return h

    def process(self, h):
        h = np.ascontiguousarray(h, dtype=np.uint64)
        me = np.zeros_like(h)
        sc = np.zeros_like(h)
        self.L.oracle_process_hist(_ptr(me), _ptr(sc), _ptr(h), h.size)
        return me, sc

    def tertiary(self, scan, is_det):
        scan = np.ascontiguousarray(scan, dtype=np.uint64)
        return self.L.oracle_tertiary(_ptr(scan), scan.size, int(is_det))

    def quaternary(self, h, scan, lmin):
        h = np.ascontiguousarray(h, dtype=np.uint64)
        scan = np.ascontiguousarray(scan, dtype=np.uint64)
        return self.L.oracle_quaternary(_ptr(h), _ptr(scan), h.size, lmin)

    def RR(self, x, tau, emb, thr, dist):
        x = np.ascontiguousarray(x)
        dt = x.dtype
        tau, emb, thr = _ints(tau), _ints(emb), _flt(thr, dt)
        out = np.zeros(tau.size * emb.size * thr.size, dtype=dt)
        err = getattr(self.L, f"oracle_RR_{self._sfx(x)}")(
            _ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size, _ptr(thr), thr.size, dist)
        return out, err

    def DET(self, x, tau, emb, lmin, thr, dist):
        return self._dl("DET", x, tau, emb, lmin, thr, dist)

    def LAM(self, x, tau, emb, vmin, thr, dist):
        return self._dl("LAM", x, tau, emb, vmin, thr, dist)

    def _dl(self, kind, x, tau, emb, lm, thr, dist):
        x = np.ascontiguousarray(x)
        dt = x.dtype
        tau, emb, lm, thr = _ints(tau), _ints(emb), _ints(lm), _flt(thr, dt)
        out = np.zeros(tau.size * emb.size * lm.size * thr.size * 5, dtype=dt)
        err = getattr(self.L, f"oracle_{kind}_{self._sfx(x)}")(
            _ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size, _ptr(lm), lm.size,
            _ptr(thr), thr.size, dist)
        return out, err


class Reference:
    """The reference CPU path compiled from its own sources (oracle/_ref)."""

    # Mangled names of the reference's public C++ API (include/AccRQA_functions.hpp).
    SYM = {
        ("RR", "f32"): "_Z9accrqa_RRPfS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
        ("RR", "f64"): "_Z9accrqa_RRPdS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
        ("DET", "f32"): "_Z10accrqa_DETPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
        ("DET", "f64"): "_Z10accrqa_DETPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
        ("LAM", "f32"): "_Z10accrqa_LAMPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
        ("LAM", "f64"): "_Z10accrqa_LAMPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    }
    PLT_CPU = 1

    @staticmethod
    def available(path=REF_SO):
        return os.path.exists(path)

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle ref`)")
        L = C.CDLL(path)
        self.L = L
        self.fn = {}
        for key, name in self.SYM.items():
            f = getattr(L, name)
            if key[0] == "RR":
                f.argtypes = [_P, _P, _SZ, _P, _I, _P, _I, _P, _I, _I, _I, _P]
            else:
                f.argtypes = [_P, _P, _SZ, _P, _I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _P]
            f.restype = None
            self.fn[key] = f
        for sfx, fl in (("f32", C.c_float), ("f64", C.c_double)):
            for kind in ("det", "lam"):
                f = getattr(L, f"ref_{kind}_hist_{sfx}")
                f.argtypes = [_P, _P, C.c_long, fl, _I, _I, _I]
                f.restype = None
            f = getattr(L, f"ref_r_element_{sfx}")
            f.argtypes = [_P, C.c_long, C.c_long, fl, _I, _I, _I]
            f.restype = _I
        for name, fl in (("ref_rp_u8_f32", C.c_float), ("ref_rp_u8_f64", C.c_double), ("ref_rp_i32_f64", C.c_double)):
            f = getattr(L, name)
            f.argtypes = [_P, _P, C.c_long, _I, _I, fl, _I]
            f.restype = _I

    @staticmethod
    def _sfx(x):
        return "f64" if np.asarray(x).dtype == np.float64 else "f32"

    def hist(self, kind, x, thr, tau, emb, dist):
        x = np.ascontiguousarray(x)
        h = np.zeros(nc_of(x.size, tau, emb) + 1, dtype=np.uint64)
        getattr(self.L, f"ref_{kind}_hist_{self._sfx(x)}")(_ptr(h), _ptr(x), x.size, thr, tau, emb, dist)
        return h

    def r_element(self, x, i, j, thr, tau, emb, dist):
        x = np.ascontiguousarray(x)
        return getattr(self.L, f"ref_r_element_{self._sfx(x)}")(_ptr(x), i, j, thr, tau, emb, dist)

    def RP(self, x, tau, emb, thr, dist, out_dtype=np.int8):
        """The reference's accrqa_RP: (N_c x N_c plot, error code); int32
        output exists for fp64 input only, as in the reference."""
        x = np.ascontiguousarray(x)
        nc = nc_of(x.size, tau, emb)
        ot = "i32" if np.dtype(out_dtype) == np.int32 else "u8"
        out = np.zeros((nc, nc), dtype=np.int32 if ot == "i32" else np.int8)
        e = getattr(self.L, f"ref_rp_{ot}_{self._sfx(x)}")(_ptr(out), _ptr(x), x.size, tau, emb, thr, dist)
        return out, e

    def RR(self, x, tau, emb, thr, dist, platform=PLT_CPU):
        x = np.ascontiguousarray(x)
        dt = x.dtype
        tau, emb, thr = _ints(tau), _ints(emb), _flt(thr, dt)
        out = np.zeros(tau.size * emb.size * thr.size, dtype=dt)
        err = C.c_int(-1)
        self.fn[("RR", self._sfx(x))](_ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size,
                                      _ptr(thr), thr.size, dist, platform, C.byref(err))
        return out, err.value

    def DET(self, x, tau, emb, lmin, thr, dist, calc_entr=1, platform=PLT_CPU):
        return self._dl("DET", x, tau, emb, lmin, thr, dist, calc_entr, platform)

    def LAM(self, x, tau, emb, vmin, thr, dist, calc_entr=1, platform=PLT_CPU):
        return self._dl("LAM", x, tau, emb, vmin, thr, dist, calc_entr, platform)

    def _dl(self, kind, x, tau, emb, lm, thr, dist, calc_entr, platform):
        x = np.ascontiguousarray(x)
        dt = x.dtype
        tau, emb, lm, thr = _ints(tau), _ints(emb), _ints(lm), _flt(thr, dt)
        out = np.zeros(tau.size * emb.size * lm.size * thr.size * 5, dtype=dt)
        err = C.c_int(-1)
        self.fn[(kind, self._sfx(x))](_ptr(out), _ptr(x), x.size, _ptr(tau), tau.size, _ptr(emb), emb.size,
                                      _ptr(lm), lm.size, _ptr(thr), thr.size, dist, calc_entr, platform,
                                      C.byref(err))
        return out, err.value


def rr_target(oracle, x, tau, emb, target_RR, dist, epsilon=0.01, max_iter=20, threshold_min=0,
              threshold_max=10):
    """Restatement of the reference's RR_target (src/accrqa/accrqa.py:470-594):
    one RR evaluation per bisection step, sequentially, through the oracle's
    RR.  Returns (threshold, RR at the last evaluated threshold, the list of
    thresholds evaluated in the loop, in order)."""
    def rr(thr):
        out, e = oracle.RR(x, [tau], [emb], np.array(thr, dtype=np.float32).astype(x.dtype), dist)
        assert e == 0
        return out.reshape(-1)

    threshold_mid = (threshold_max - threshold_min) / 2.0 + threshold_min
    v = rr([threshold_min, threshold_mid, threshold_max])
    low_RR, hgh_RR = v[0], v[2]
    seen = []
    current_RR = None
    for _ in range(1, max_iter + 1):
        threshold_mid = (threshold_max - threshold_min) / 2.0 + threshold_min
        seen.append(threshold_mid)
        current_RR = rr([threshold_mid])[0]
        if abs(current_RR - target_RR) < epsilon:
            return threshold_mid, current_RR, seen
        if (hgh_RR - low_RR) < epsilon:
            return threshold_mid, current_RR, seen
        if current_RR < target_RR:
            threshold_min = threshold_mid
        else:
            threshold_max = threshold_mid
    threshold_mid = (threshold_max - threshold_min) / 2.0 + threshold_min
    return threshold_mid, current_RR, seen

