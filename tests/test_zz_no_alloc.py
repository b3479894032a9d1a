"""The no-allocation entry point (accrqa_amd_rqa_device_static_f32) against
the entry point it is a variant of (accrqa_amd_rqa_device_f32), through the
C ABI: RR bit for bit, DET/LAM measures to the parity bar (ENTR within 1e-6
relative), on two calls in a row, the second on a smaller input, in both
norms, with a threshold that never recurs and two l_min values; and sizes
beyond its fixed work memory rejected with ERR_INVALID_ARGUMENT."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_measures, same_bits, say

pytestmark = pytest.mark.gpu


def _run(amd, torch, fn, dx, n, tau, emb, thr, lm, dist):
    T, nl = len(thr), len(lm)
    o = [torch.full((k,), -7.0, dtype=torch.float32, device="cuda") for k in (T, T * nl * 5, nl * T * 5)]
    t = np.ascontiguousarray(thr, dtype=np.float32)
    l = np.ascontiguousarray(lm, dtype=np.int32)
    s = torch.cuda.current_stream()
    e = fn(C.c_void_p(dx.data_ptr()), n, tau, emb, t.ctypes.data, T, l.ctypes.data, nl, dist,
           C.c_void_p(o[0].data_ptr()), C.c_void_p(o[1].data_ptr()), C.c_void_p(o[2].data_ptr()),
           C.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    return e, [v.cpu().numpy() for v in o]


def test_static_entry_matches_device_entry():
    import torch
    import accrqa_amd as amd
    L = amd.load()
    fs = L.accrqa_amd_rqa_device_static_f32
    fd = L.accrqa_amd_rqa_device_f32   # both bound by accrqa_amd.load()
    rng = np.random.default_rng(123)
    w = np.cumsum(rng.standard_normal(32768))
    w = ((w - w.mean()) / w.std()).astype(np.float32)
    cases = [(w, 2, 5, [0.05, 0.2, -1.0, 0.5, 0.1], [2, 3]),            # full size limit first
             (w[:3001].copy(), 1, 3, [0.3, 0.05, 0.8], [2, 1]),           # then a smaller input
             (w[:777].copy(), 3, 2, list(np.linspace(0.02, 0.9, 20)), [2, 4])]
    for dist in (1, 2):
        for x, tau, emb, thr, lm in cases:
            say("static", dist, x.size, tau, emb, len(thr))
            dx = torch.from_numpy(x).cuda()
            e1, a = _run(amd, torch, fs, dx, x.size, tau, emb, thr, lm, dist)
            e2, b = _run(amd, torch, fd, dx, x.size, tau, emb, thr, lm, dist)
            assert e1 == 0 and e2 == 0, (e1, e2)
            assert same_bits(a[0], b[0]), (dist, x.size, "RR", a[0], b[0])
            assert_measures(a[1], b[1], np.float32, f"static DET {dist} {x.size}")
            assert_measures(a[2], b[2], np.float32, f"static LAM {dist} {x.size}")
    # beyond the fixed work memory: rejected, outputs untouched
    big = torch.zeros(32769, dtype=torch.float32, device="cuda")
    e, out = _run(amd, torch, fs, big, 32769, 1, 2, [0.1], [2], 1)
    assert e == 2 and np.all(out[0] == -7.0)
    e, out = _run(amd, torch, fs, dx, x.size, 1, 2, list(np.linspace(0.1, 1, 33)), [2], 1)
    assert e == 2
