"""Parity at the full sizes of the benchmark configurations.

An O(N_c^2) CPU oracle run at N = 262,144 or 1,048,576 takes far longer than a
test may, so these tests check what the domain fixes independently of size:

* the MFMA form and the exact form of the Euclidean decision agree bit for bit
  on every histogram bin (two different kernels and arithmetics);
* the split form (rank tiles kept, LAM row runs and DET read from them) and
  the direct form agree bit for bit;
* for every threshold, sum_l l*h_V[l] (LAM pass) = sum_l l*h_D[l] (DET pass,
  doubled, with the line of identity) = the recurrence count of the separate
  RR kernel (k_rr), all three from different kernels.

The oracle itself is pinned at small sizes (tests/test_oracle.py) and the GPU
against it there (tests/test_gpu_parity.py); these tests carry that parity to
the configurations' real shapes.
"""
import numpy as np
import pytest

from accrqa_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[fullsize]", *a, flush=True)


def hists(amd, x, tau, emb, thr, dist, which=7):
    return amd.histograms(x, tau, emb, np.asarray(thr, dtype=x.dtype), dist, which=which)


def assert_same(a, b, what):
    for key in ("det", "lam"):
        if key in a and key in b:
            bad = np.nonzero((a[key] != b[key]).any(axis=1))[0]
            assert bad.size == 0, (what, key, "thresholds differ:", bad[:8])


def check_counts(amd, x, tau, emb, thr, dist, h, what):
    rr = hists(amd, x, tau, emb, thr, dist, which=1)["rr"]
    nc = h["lam"].shape[1] - 1
    lengths = np.arange(nc + 1, dtype=np.uint64)
    for k in range(len(thr)):
        lam = int((lengths * h["lam"][k]).sum())
        det = int((lengths * h["det"][k]).sum())
        assert lam == int(rr[k]), (what, k, "LAM", lam, int(rr[k]))
        assert det == int(rr[k]), (what, k, "DET", det, int(rr[k]))


@pytest.fixture
def forms():
    import accrqa_amd as amd
    amd.load()
    yield amd
    amd.set_form(amd.FORM_AUTO)
    amd.set_split(amd.SPLIT_AUTO)


def test_config3_gate_shape(forms):
    """AR(1)+sine, N = 262,144, m = 16, 16 thresholds (the bench workload)."""
    amd = forms
    c = W.config3(m=16)
    x, tau, emb, thr = c["x"], c["tau"][0], c["emb"][0], c["thr"]
    amd.set_form(amd.FORM_EXACT)
    say("config3 exact split")
    ex = hists(amd, x, tau, emb, thr, 1)
    amd.set_form(amd.FORM_MFMA)
    say("config3 mfma split")
    mx = hists(amd, x, tau, emb, thr, 1)
    assert_same(ex, mx, "config3 exact vs mfma")
    amd.set_split(amd.SPLIT_NEVER)
    say("config3 mfma direct")
    md = hists(amd, x, tau, emb, thr, 1)
    assert_same(mx, md, "config3 split vs direct")
    amd.set_split(amd.SPLIT_AUTO)
    say("config3 counts")
    check_counts(amd, x, tau, emb, thr, 1, mx, "config3")


def test_config3_fp64_input(forms):
    """The gate workload in fp64: the MFMA form (hi/lo Gram of the series
    rounded to fp32, exact re-decision in fp64) against the exact fp64 form."""
    amd = forms
    c = W.config3(m=16, dtype=np.float64)
    x, tau, emb, thr = c["x"], c["tau"][0], c["emb"][0], c["thr"]
    amd.set_form(amd.FORM_EXACT)
    say("config3 fp64 exact")
    ex = hists(amd, x, tau, emb, thr, 1)
    amd.set_form(amd.FORM_MFMA)
    say("config3 fp64 mfma")
    mx = hists(amd, x, tau, emb, thr, 1)
    assert_same(ex, mx, "config3 fp64 exact vs mfma")
    say("config3 fp64 counts")
    check_counts(amd, x, tau, emb, thr, 1, mx, "config3 fp64")


def test_config2_both_norms(forms):
    """Lorenz x, N = 131,072, tau = 8, m = 3, 20 thresholds, both norms."""
    amd = forms
    c = W.config2()
    x, tau, emb, thr = c["x"], c["tau"][0], c["emb"][0], c["thr"]
    for dist in (1, 2):
        say("config2 split", dist)
        sp = hists(amd, x, tau, emb, thr, dist)
        amd.set_split(amd.SPLIT_NEVER)
        say("config2 direct", dist)
        di = hists(amd, x, tau, emb, thr, dist)
        amd.set_split(amd.SPLIT_AUTO)
        assert_same(sp, di, f"config2 dist {dist} split vs direct")
        if dist == 1:
            amd.set_form(amd.FORM_MFMA)
            say("config2 mfma")
            mx = hists(amd, x, tau, emb, thr, dist)
            amd.set_form(amd.FORM_AUTO)
            assert_same(sp, mx, "config2 exact vs mfma")
        check_counts(amd, x, tau, emb, thr, dist, sp, f"config2 dist {dist}")


def test_config5_million(forms):
    """Rossler x, N = 1,048,576, m = 10, 32 thresholds: the rank tiles
    (550 GB) do not fit whole, so the split form runs slab by slab; both
    distance forms, and the direct form (DET recomputing its pairs), agree
    bit for bit."""
    amd = forms
    c = W.config5()
    x, tau, emb, thr = c["x"], c["tau"][0], c["emb"][0], c["thr"]
    amd.set_form(amd.FORM_MFMA)
    say("config5 mfma (slab form)")
    mx = hists(amd, x, tau, emb, thr, 1)
    amd.set_form(amd.FORM_EXACT)
    say("config5 exact (slab form)")
    ex = hists(amd, x, tau, emb, thr, 1)
    amd.set_form(amd.FORM_AUTO)
    assert_same(ex, mx, "config5 exact vs mfma")
    try:
        amd.set_split(amd.SPLIT_NEVER)
        say("config5 direct form")
        dr = hists(amd, x, tau, emb, thr, 1)
    finally:
        amd.set_split(amd.SPLIT_AUTO)
    assert_same(dr, mx, "config5 direct vs slab")
    say("config5 counts")
    check_counts(amd, x, tau, emb, thr, 1, mx, "config5")


def test_rp_large(forms):
    """The recurrence plot at N_c = 24,000 (576 MB) on the default workload's
    series: symmetric, ones on the diagonal, and its total equals the
    recurrence count of the separate RR kernel for the same threshold, in
    both norms."""
    import torch
    amd = forms
    c = W.config3(m=16)
    x = np.ascontiguousarray(c["x"][:24015])
    thr = float(c["thr"][5])
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    nc = x.size - 15
    dout = torch.empty(nc * nc, dtype=torch.int8, device=dev)
    for dist in (1, 2):
        say("rp", nc, dist)
        amd.rp_device(dx.data_ptr(), x.size, 1, 16, thr, dist, dout.data_ptr())
        torch.cuda.synchronize()
        rp = dout.view(nc, nc)
        assert torch.equal(rp, rp.t())
        assert bool((torch.diagonal(rp) == 1).all())
        total = int(rp.sum(dtype=torch.int64).item())
        rr = hists(amd, x, 1, 16, [thr], dist, which=1)["rr"]
        assert total == int(rr[0]), (dist, total, int(rr[0]))
    del dout
    torch.cuda.empty_cache()
