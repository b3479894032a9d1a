"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden fixtures.

Bar (from the build's contract): recurrence counts and line-length histograms
bit-exact; RR, DET/LAM, L/TT and Lmax/TTmax bit-exact (they are integer ratios
divided exactly as the reference divides); ENTR within 1e-6 relative for fp32
outputs and 1e-12 relative for fp64 outputs (the GPU sums p*log(p) in a
different order with the device log).
"""
import numpy as np
import pytest

from conftest import assert_measures, case_arrays, check_case, dense, same_bits, say, unhex

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import accrqa_amd
    accrqa_amd.load()
    return accrqa_amd


# ----------------------------------------------------------------- golden fixtures


def test_golden_histograms_bit_exact(amd, golden):
    for case in golden["cases"]:
        say("golden hist", case["name"])
        dt, x, _ = case_arrays(case)
        by_key = {}
        for h in case["hist"]:
            by_key.setdefault((h["tau"], h["emb"]), []).append(h)
        for (tau, emb), hs in by_key.items():
            thr = np.array([float.fromhex(h["thr"]) for h in hs], dtype=dt)
            got = amd.histograms(x, tau, emb, thr, case["dist"], which=7)
            for t, h in enumerate(hs):
                hd = dense(h["hD"], h["nc"])
                hv = dense(h["hV"], h["nc"])
                assert np.array_equal(got["det"][t], hd), (case["name"], tau, emb, h["thr"], "h_D")
                assert np.array_equal(got["lam"][t], hv), (case["name"], tau, emb, h["thr"], "h_V")
                lv = np.arange(h["nc"] + 1, dtype=np.uint64)
                assert int(got["rr"][t]) == int((lv * hv).sum()), (case["name"], "count")


def test_golden_api_outputs(amd, golden):
    for case in golden["cases"]:
        say("golden api", case["name"])
        dt, x, thr = case_arrays(case)
        d = case["dist"]
        rr = amd.RR(x, case["tau"], case["emb"], thr, d)
        assert same_bits(rr.reshape(-1), unhex(case["RR"], dt)), case["name"]
        det = amd.DET(x, case["tau"], case["emb"], case["lmin"], thr, d)
        assert_measures(det, unhex(case["DET"], dt), dt, case["name"] + " DET")
        lam = amd.LAM(x, case["tau"], case["emb"], case["lmin"], thr, d)
        assert_measures(lam, unhex(case["LAM"], dt), dt, case["name"] + " LAM")


def test_crqa_known_answers_on_gpu(amd, crqa_kat, crqa_series):
    for rec in crqa_kat["DET"]:
        d = 1 if rec["norm"] == "euclidean" else 2
        out = amd.DET(crqa_series, [int(rec["tau"])], [int(rec["emb"])], [int(rec["lmin"])], [rec["thr"]], d)
        v = out.reshape(5)
        assert abs(v[0] - rec["CRQA_DET"]) <= 1e-4 and abs(v[1] - rec["CRQA_L"]) <= 1e-4
        assert abs(v[3] - rec["CRQA_ENTR"]) <= 1e-4 and abs(v[4] - rec["CRQA_RR"]) <= 1e-4
    for rec in crqa_kat["LAM"]:
        d = 1 if rec["norm"] == "euclidean" else 2
        out = amd.LAM(crqa_series, [int(rec["tau"])], [int(rec["emb"])], [int(rec["vmin"])], [rec["thr"]], d)
        v = out.reshape(5)
        assert abs(v[0] - rec["CRQA_LAM"]) <= 1e-4 and abs(v[1] - rec["CRQA_TT"]) <= 1e-4
        assert abs(v[4] - rec["CRQA_RR"]) <= 1e-4


# ----------------------------------------------------------------- oracle, seeded inputs


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dist", [1, 2])
def test_multi_segment_multi_group(amd, oracle, dt, dist):
    """N_c of a few thousand: several row/diagonal segments (stitch path), and
    more thresholds than one group (group planner + permutation)."""
    rng = np.random.default_rng(100 + dist)
    x = np.cumsum(rng.standard_normal(4500)).astype(dt)
    x = (x - x.mean()) / x.std()
    thrs = sorted(rng.uniform(0.01, 0.6, size=37).astype(dt))[::-1]   # unsorted order on purpose
    say("multi", dt, dist)
    check_case(amd, oracle, x.astype(dt), [1, 3], [2, 5], [2, 7], thrs, dist, f"multi {dt} {dist}")


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_random_shapes(amd, oracle, dt):
    rng = np.random.default_rng(7)
    for trial in range(5):
        n = int(rng.integers(50, 2600))
        x = rng.standard_normal(n).astype(dt)
        tau = int(rng.integers(1, 9))
        emb = int(rng.integers(1, 12))
        if n - (emb - 1) * tau < 2:
            continue
        thrs = list(rng.uniform(0.1, 4.0, size=int(rng.integers(1, 9))).astype(dt))
        dist = 1 + trial % 2
        say("random", trial, n, tau, emb, len(thrs), dist)
        check_case(amd, oracle, x, [tau], [emb], [1, 2, 3], thrs, dist, f"random {trial}")


def test_long_lines_beyond_lds_cap_and_segments(amd, oracle):
    """Runs longer than the LDS histogram cap (256) and than a segment (1024
    steps): plateaus of 1,300 nearly equal samples give long diagonal and
    vertical lines that cross segment borders."""
    rng = np.random.default_rng(4)
    levels = [0.0, 1.0, 0.0, 2.0, 1.0]
    x = np.concatenate([np.full(1300, v) for v in levels]) + 1e-3 * rng.standard_normal(6500)
    x = x.astype(np.float32)
    for dist in (1, 2):
        say("long lines", dist)
        got = amd.histograms(x, 1, 2, np.array([0.05, 0.3], np.float32), dist, which=7)
        for k, thr in enumerate([0.05, 0.3]):
            hd = oracle.hist("det", x, np.float32(thr), 1, 2, dist)
            hv = oracle.hist("lam", x, np.float32(thr), 1, 2, dist)
            assert np.array_equal(got["det"][k], hd)
            assert np.array_equal(got["lam"][k], hv)
            assert hv[1025:].sum() > 0 and hd[1025:-1].sum() > 0   # the case really has long lines


def test_constant_series_full_recurrence(amd):
    for dt in (np.float32, np.float64):
        for n in (1, 2, 3, 64, 65, 1500):
            x = np.full(n, 3.25, dtype=dt)
            say("constant", dt, n)
            got = amd.histograms(x, 1, 1, np.array([0.0, 1.0], dtype=dt), 1, which=7)
            for k in range(2):
                assert int(got["rr"][k]) == n * n
                hv = np.zeros(n + 1, np.uint64)
                hv[n] = n
                assert np.array_equal(got["lam"][k], hv)
                hd = np.zeros(n + 1, np.uint64)
                for r in range(1, n):
                    hd[n - r] += 2
                hd[n] += 1
                assert np.array_equal(got["det"][k], hd)


def test_special_values(amd, oracle):
    rng = np.random.default_rng(9)
    for dt in (np.float32, np.float64):
        x = rng.standard_normal(700).astype(dt)
        x[[5, 300]] = np.nan
        x[600] = np.copysign(np.nan, -1.0)   # a NaN with the sign bit set
        assert np.signbit(x[600])
        x[[100, 101, 102]] = np.inf
        x[400] = -np.inf
        x[500:520] = 0.0
        x[520] = np.finfo(dt).tiny / 4   # subnormal differences
        thrs = [-1.0, -0.0, 0.0, np.finfo(dt).tiny, 1e-30, 0.7, np.inf, np.nan]
        for dist in (1, 2):
            say("special", dt, dist)
            check_case(amd, oracle, x, [1, 2], [1, 3], [1, 2], thrs, dist, f"special {dt} {dist}")


def test_tie_thresholds_exact(amd, oracle):
    """Integer data: squared distances are exact integers and thresholds sit
    exactly on attainable distances (sqrt(k) rounded), the worst case for the
    squared-bound rewrite of the square root."""
    rng = np.random.default_rng(13)
    for dt in (np.float32, np.float64):
        x = rng.integers(-6, 7, size=1300).astype(dt)
        thrs = [dt(np.sqrt(dt(k))) for k in (1, 2, 3, 5, 8, 13, 50, 99)]
        for dist in (1, 2):
            say("ties", dt, dist)
            check_case(amd, oracle, x, [1, 4], [3, 6], [2], thrs, dist, f"ties {dt} {dist}")


def test_many_thresholds_with_never_recurring(amd, oracle):
    """More thresholds than one group holds (100: three recurring groups) and
    more never-recurring thresholds (70, zero or negative) than the count
    buffer's default size, shuffled: measures and raw histograms per threshold,
    in the caller's order."""
    rng = np.random.default_rng(21)
    for dt in (np.float32, np.float64):
        x = rng.standard_normal(700).astype(dt)
        thrs = np.concatenate([-rng.random(69), [0.0], np.linspace(0.05, 2.5, 30)]).astype(dt)
        thrs = thrs[rng.permutation(thrs.size)]
        thrs = np.concatenate([thrs, thrs[:5]])   # duplicates across groups
        for dist in (1, 2):
            say("many thresholds", dt, dist)
            check_case(amd, oracle, x, [2], [3], [2], list(thrs), dist, f"many {dt} {dist}")
            h = amd.histograms(x, 2, 3, thrs, dist, which=7)
            for k, t in enumerate(thrs):
                assert int(h["rr"][k]) == oracle.rr_count(x, t, 2, 3, dist), (dt, dist, k, t)
                assert np.array_equal(h["det"][k], oracle.hist("det", x, t, 2, 3, dist)), (dt, dist, k, t)
                assert np.array_equal(h["lam"][k], oracle.hist("lam", x, t, 2, 3, dist)), (dt, dist, k, t)


def test_argument_errors_on_gpu(amd):
    x = np.ones(10, dtype=np.float32)
    with pytest.raises(amd.AccrqaError) as ei:
        amd.RR(x, [4], [4], [0.5], 1)        # N_c < 1
    assert ei.value.code == 2
    with pytest.raises(amd.AccrqaError) as ei:
        amd.DET(x, [1], [1], [11], [0.5], 1)   # l_min > N_c
    assert ei.value.code == 2
    with pytest.raises(amd.AccrqaError) as ei:
        amd.LAM(x, [1], [1], [2], [0.5], 0)    # invalid distance
    assert ei.value.code == 2


def test_cxx_api_platform_dispatch(amd, oracle):
    """The reference's mangled C++ entry point runs on the GPU with PLT_AMD_GPU
    and with every other platform value (the reference's fallback to its
    default platform), giving the oracle's results."""
    import ctypes as C
    L = amd.load()
    f = getattr(L, "_Z10accrqa_DETPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error")
    P, I = C.c_void_p, C.c_int
    f.argtypes = [P, P, C.c_size_t, P, I, P, I, P, I, P, I, I, I, I, P]
    f.restype = None
    rng = np.random.default_rng(21)
    x = rng.standard_normal(900).astype(np.float32)
    tau = np.array([2], np.int32)
    emb = np.array([3], np.int32)
    lm = np.array([2, 3], np.int32)
    thr = np.array([0.5, 1.0], np.float32)
    out = np.zeros(2 * 2 * 5, np.float32)
    err = C.c_int(-1)
    args = [out.ctypes.data, x.ctypes.data, x.size, tau.ctypes.data, 1, emb.ctypes.data, 1,
            lm.ctypes.data, 2, thr.ctypes.data, 2, 1, 1]
    f(*args, 2048, C.byref(err))
    assert err.value == 0
    exp, _ = oracle.DET(x, [2], [3], [2, 3], [0.5, 1.0], 1)
    assert_measures(out, exp, np.float32, "cxx DET")
    # PLT_CPU, PLT_NV_GPU and an unknown value are served on the GPU, as the
    # reference serves platforms it lacks on its default one
    for plt in (1, 1024, 7):
        out[:] = 0
        err.value = -1
        f(*args, plt, C.byref(err))
        assert err.value == 0, plt
        assert_measures(out, exp, np.float32, f"cxx DET platform {plt}")


def test_device_entry_point_matches_host(amd, oracle):
    import torch
    rng = np.random.default_rng(31)
    x = rng.standard_normal(3000).astype(np.float32)
    thr = [0.3, 0.6, 0.9]
    lm = [2, 4]
    dx = torch.from_numpy(x).cuda()
    d_rr = torch.zeros(3, dtype=torch.float32, device="cuda")
    d_det = torch.zeros(3 * 2 * 5, dtype=torch.float32, device="cuda")
    d_lam = torch.zeros(2 * 3 * 5, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream()
    say("device entry")
    amd.rqa_device(dx.data_ptr(), x.size, 2, 4, thr, lm, 1, d_rr.data_ptr(), d_det.data_ptr(), d_lam.data_ptr(),
                   stream=s.cuda_stream, dtype=np.float32)
    torch.cuda.synchronize()
    rr_o, _ = oracle.RR(x, [2], [4], thr, 1)
    assert same_bits(d_rr.cpu().numpy(), rr_o)
    det_o, _ = oracle.DET(x, [2], [4], lm, thr, 1)
    assert_measures(d_det.cpu().numpy(), det_o, np.float32, "device DET")
    lam_o, _ = oracle.LAM(x, [2], [4], lm, thr, 1)
    assert_measures(d_lam.cpu().numpy(), lam_o, np.float32, "device LAM")


def test_threshold_sets_for_rank_table(amd, oracle):
    """Threshold lists that stress the bucket table of the rank computation:
    exact duplicates, neighbours one ulp apart (several bounds in one bucket),
    a span of 60 decades, never-recurring values mixed in, and more than 32
    thresholds (several groups)."""
    rng = np.random.default_rng(57)
    for dt in (np.float32, np.float64):
        x = np.cumsum(rng.standard_normal(1800)).astype(dt)
        x = ((x - x.mean()) / x.std()).astype(dt)
        a = dt(0.3)
        near = [a, np.nextafter(a, dt(1)), np.nextafter(np.nextafter(a, dt(1)), dt(1)), a, dt(0.3000001)]
        wide = [dt(1e-30), dt(1e-10), dt(1e-3), dt(0.05), dt(2.0), dt(1e10), dt(1e30)]
        dead = [dt(-2.0), dt(np.nan), dt(-0.0)]
        many = list(np.linspace(0.01, 1.5, 41).astype(dt))
        for name, thrs in (("near", near), ("wide", wide), ("mixed", near + dead + wide), ("many", many)):
            for dist in (1, 2):
                say("rank table", dt, name, dist)
                check_case(amd, oracle, x, [1, 3], [2], [2], thrs, dist, f"rank {name} {dt} {dist}")


@pytest.fixture
def mfma_form(amd):
    amd.set_form(amd.FORM_MFMA)
    yield amd
    amd.set_form(amd.FORM_AUTO)


def test_mfma_form_bit_exact(mfma_form, oracle):
    """The MFMA form (Gram on matrix cores + exact recheck near thresholds)
    against the oracle: every decision, histogram and measure as the exact
    form's.  fp32 Euclidean cases of several shapes, ties and special values."""
    amd = mfma_form
    rng = np.random.default_rng(77)
    cases = []
    x = np.cumsum(rng.standard_normal(3500)).astype(np.float32)
    cases.append(("walk", ((x - x.mean()) / x.std()).astype(np.float32), [1, 2], [16, 5], [2, 3],
                  list(np.linspace(0.2, 3.0, 21).astype(np.float32))))
    cases.append(("ties", rng.integers(-4, 5, size=1300).astype(np.float32), [1, 3], [8, 3], [2],
                  [np.float32(np.sqrt(np.float32(k))) for k in (1, 4, 9, 17, 40, 61)]))
    xs = rng.standard_normal(900).astype(np.float32)
    xs[[10, 500]] = np.nan
    xs[300] = np.inf
    cases.append(("special", xs, [1], [9], [2], [-1.0, 0.0, 0.5, 2.0, np.inf]))
    cases.append(("m33", rng.standard_normal(2100).astype(np.float32), [2], [33], [2, 4],
                  list(np.linspace(1.0, 12.0, 7).astype(np.float32))))
    # tau = 0 (every coordinate the same sample): dimensions past m must add
    # nothing although their window offsets coincide with real ones; odd m,
    # unused steps in the first chunk of 16 dimensions, and a second chunk
    cases.append(("tau0", rng.standard_normal(1500).astype(np.float32), [0, 3], [10, 21, 5], [2],
                  list(np.linspace(0.1, 4.0, 9).astype(np.float32))))
    # the Gram's scale and split: a spike 10^6 times the rest (the other
    # samples' low halves are fp16 subnormals), and series of tiny and huge
    # magnitude (scale 2^e at both clamps' sides)
    xs = (1e-3 * rng.standard_normal(1700)).astype(np.float32)
    xs[777] = 1e3
    cases.append(("spike", xs, [1], [16, 9], [2], list(np.linspace(2e-4, 8e-3, 11).astype(np.float32))))
    cases.append(("tiny", (1e-18 * rng.standard_normal(1200)).astype(np.float32), [1], [16], [2],
                  list(np.linspace(1e-19, 6e-18, 9).astype(np.float32))))
    cases.append(("huge", (1e12 * rng.standard_normal(1200)).astype(np.float32), [2], [16], [2],
                  list(np.linspace(1e11, 6e12, 9).astype(np.float32))))
    # fp64 input: hi/lo Gram of the series rounded to fp32, exact re-decision in fp64
    # (the same cases, the thresholds and series widened, plus one series
    # whose fp64 values are not fp32 values)
    cases += [(n + " fp64", x.astype(np.float64), taus, embs, lm, [float(t) for t in thrs])
              for n, x, taus, embs, lm, thrs in list(cases)]
    w = np.cumsum(rng.standard_normal(2500))
    cases.append(("walk fp64 fine", (w - w.mean()) / w.std(), [1], [12], [2],
                  list(np.linspace(0.2, 3.0, 17))))
    for name, x, taus, embs, lm, thrs in cases:
        say("mfma", name)
        dt = x.dtype.type
        check_case(amd, oracle, x, taus, embs, lm, thrs, 1, f"mfma {name}")
        got = amd.histograms(x, taus[0], embs[0], np.asarray(thrs, dt), 1, which=7)
        for k, t in enumerate(thrs):
            assert np.array_equal(got["lam"][k], oracle.hist("lam", x, dt(t), taus[0], embs[0], 1)), (name, t)
            assert np.array_equal(got["det"][k], oracle.hist("det", x, dt(t), taus[0], embs[0], 1)), (name, t)


def test_mfma16_instruction_properties(amd):
    """What the MFMA form's error bound takes of v_mfma_f32_32x32x16_f16, in
    the Gram kernels' lane layout: products exact, fp16 subnormal operands
    kept, and the result within 16 additions of relative error 2^-23 each of
    the exact c + a b against the operands' nominal sizes, i.e.
    |d - (c + ab)| <= 32u (|c| + |a~||b~|), u = 2^-24, where a~ counts an fp16
    subnormal at 2^-14, the smallest normal (the instruction aligns such a
    product by that exponent: against |a||b| alone the error reaches 10^3 u
    when subnormal and normal operands meet in one sum).  A layout fault (a
    and b in different dimension slots) fails every case."""
    rng = np.random.default_rng(3)
    u = 2.0 ** -24
    g16 = 32 * u / (1 - 32 * u)

    def check(a, b, c, what, exact_result=False):
        a = np.asarray(a, np.float16)
        b = np.asarray(b, np.float16)
        c = np.asarray(c, np.float32)
        d = amd.mfma16_probe(a, b, c).astype(np.float64)
        A, B, Cc = a.astype(np.float64), b.astype(np.float64), c.astype(np.float64)
        want = Cc + A @ B   # fp16 products are exact in fp64; 17-term sums off by < 2^-48 relative
        if exact_result:
            assert np.array_equal(d, want), (what, np.abs(d - want).max())
        nA = np.where(A == 0, 0.0, np.maximum(np.abs(A), 2.0 ** -14))
        nB = np.where(B == 0, 0.0, np.maximum(np.abs(B), 2.0 ** -14))
        bound = g16 * (np.abs(Cc) + nA @ nB) * (1 + 2.0 ** -40)
        assert np.all(np.abs(d - want) <= bound), (what, (np.abs(d - want) / np.maximum(bound, 1e-300)).max())

    z = np.zeros((32, 32), np.float32)
    # layout and plain sums
    check(rng.standard_normal((32, 16)), rng.standard_normal((16, 32)), z, "normal")
    check(rng.integers(-8, 9, (32, 16)), rng.integers(-8, 9, (16, 32)), rng.integers(-50, 50, (32, 32)),
          "integers", exact_result=True)
    # exponents spread over the fp16 range, random signs, and an accumulator
    for s in range(4):
        ea, eb = rng.integers(-14, 15, (32, 16)), rng.integers(-14, 15, (16, 32))
        a = rng.choice([-1.0, 1.0], (32, 16)) * (1 + rng.random((32, 16))) * 2.0 ** ea
        b = rng.choice([-1.0, 1.0], (16, 32)) * (1 + rng.random((16, 32))) * 2.0 ** eb
        c = (rng.standard_normal((32, 32)) * 2.0 ** rng.integers(-20, 20, (32, 32))).astype(np.float32)
        check(a, b, c, f"spread {s}")
    # cancellation: pairs of equal and opposite products plus small terms
    a = rng.standard_normal((32, 16)) * 100
    a[:, 1::2] = a[:, 0::2]
    b = rng.standard_normal((16, 32)) * 100
    b[1::2, :] = -b[0::2, :] * (1 + 2.0 ** -10 * rng.integers(-3, 4, (8, 32)))
    check(a, b, z, "cancel")
    # fp16 subnormal operands (2^-24 .. 2^-15) against normal ones: kept, not flushed
    sub = rng.integers(1, 1024, (32, 16)) * 2.0 ** -24
    check(sub, np.full((16, 32), 2.0 ** 14), z, "subnormal x 2^14", exact_result=True)
    check(rng.standard_normal((32, 16)) * 2.0 ** 13, rng.integers(-1023, 1024, (16, 32)) * 2.0 ** -24, z,
          "normal x subnormal")
    assert np.all(amd.mfma16_probe(np.full((32, 16), 2.0 ** -24, np.float16), np.ones((16, 32), np.float16), z)
                  == np.float32(16 * 2.0 ** -24))
    # subnormal and normal operands in one sum, with and without an accumulator
    # (the hi/lo Gram's l halves of small samples are such operands)
    for s in range(4):
        ea, eb = rng.integers(-30, 13, (32, 16)), rng.integers(-30, 13, (16, 32))
        a = rng.choice([-1.0, 1.0], (32, 16)) * (1 + rng.random((32, 16))) * 2.0 ** ea
        b = rng.choice([-1.0, 1.0], (16, 32)) * (1 + rng.random((16, 32))) * 2.0 ** eb
        c = z if s < 2 else (rng.standard_normal((32, 32)) * 2.0 ** rng.integers(-10, 10, (32, 32))).astype(np.float32)
        check(a, b, c, f"mixed subnormal {s}")


@pytest.mark.gpu
def test_split_and_direct_forms(amd, oracle):
    """The split form (upper-triangle rank tiles kept, LAM row runs and DET
    read from them) and the direct form (full-column LAM, DET recomputed)
    give the oracle's histograms bit for bit: both norms, fp32 and fp64,
    sequence lengths around the 64-column band and 32-step block sizes, and
    series long enough for many segments per column."""
    rng = np.random.default_rng(2024)
    cases = []
    for n, m, tau in ((40, 3, 1), (64 + 2, 3, 1), (129, 2, 2), (700, 4, 3), (3001, 5, 1), (4200, 10, 2)):
        w = np.cumsum(rng.standard_normal(n))
        cases.append((f"walk{n}", (w - w.mean()) / w.std(), m, tau))
    try:
        for name, x64, m, tau in cases:
            for dt in (np.float32, np.float64):
                x = x64.astype(dt)
                thr = np.array([0.05, 0.3, 0.9, 2.0], dt)
                for dist in (1, 2):
                    say("split", name, dt.__name__, dist)
                    amd.set_split(amd.SPLIT_AUTO)
                    a = amd.histograms(x, tau, m, thr, dist, which=7)
                    amd.set_split(amd.SPLIT_NEVER)
                    b = amd.histograms(x, tau, m, thr, dist, which=7)
                    for key in ("rr", "det", "lam"):
                        assert np.array_equal(a[key], b[key]), (name, dt, dist, key)
                    for k in (0, 2):
                        for kind in ("lam", "det"):
                            want = oracle.hist(kind, x, dt(thr[k]), tau, m, dist)
                            assert np.array_equal(a[kind][k], want), (name, dt, dist, kind, thr[k])
    finally:
        amd.set_split(amd.SPLIT_AUTO)


def test_overlap_identical(amd, oracle):
    """The split form's DET scan beside the LAM scans on a second stream
    (accrqa_amd_set_overlap(1), the default) against the same scans one after
    the other: every histogram and measure is the same, and equals the
    oracle's, for the whole tile store and slab by slab, the MFMA and exact
    forms, both norms, fp32 and fp64, one and two threshold groups, and DET or
    LAM asked alone; several calls in a row keep the order."""
    rng = np.random.default_rng(5150)
    w = np.cumsum(rng.standard_normal(3001))
    x64 = (w - w.mean()) / w.std()
    thr16 = np.linspace(0.1, 2.0, 16)
    thr21 = np.linspace(0.05, 2.5, 21)
    try:
        for slab in (0, 512):
            amd.set_slab_rows(slab)
            for form in (amd.FORM_AUTO, amd.FORM_EXACT):
                amd.set_form(form)
                for dt in (np.float32, np.float64):
                    x = x64.astype(dt)
                    for dist in (1, 2):
                        for thr in (thr16.astype(dt), thr21.astype(dt)):
                            say("overlap", slab, form, dt.__name__, dist, thr.size)
                            got = {}
                            for on in (0, 1, 1):
                                amd.set_overlap(on)
                                got[on] = (amd.histograms(x, 1, 10, thr, dist, which=7),
                                           amd.DET(x, [1], [10], [2, 3], thr, dist),
                                           amd.LAM(x, [1], [10], [2, 3], thr, dist),
                                           amd.histograms(x, 1, 10, thr, dist, which=2))
                            a, b = got[0], got[1]
                            for key in ("rr", "det", "lam"):
                                assert np.array_equal(a[0][key], b[0][key]), (slab, form, dt, dist, thr.size, key)
                            assert same_bits(a[1], b[1]) and same_bits(a[2], b[2]), (slab, form, dt, dist, thr.size)
                            assert np.array_equal(a[3]["det"], b[3]["det"]), (slab, form, dt, dist, "DET alone")
                            if slab == 0 and dt == np.float32 and thr.size == 16:
                                for k, t in enumerate(thr):
                                    assert np.array_equal(b[0]["det"][k], oracle.hist("det", x, t, 1, 10, dist)), (form, dist, t)
                                    assert np.array_equal(b[0]["lam"][k], oracle.hist("lam", x, t, 1, 10, dist)), (form, dist, t)
        assert amd.load().accrqa_amd_set_overlap(2) != 0 and amd.load().accrqa_amd_set_overlap(-1) != 0
    finally:
        amd.set_overlap(1)
        amd.set_slab_rows(0)
        amd.set_form(amd.FORM_AUTO)


def test_scan_waves_identical(amd, oracle):
    """The tile scans held to 7 (the default), 4 and 1 waves per SIMD, and
    uncapped (accrqa_amd_set_scan_waves): every histogram is the same, and
    equals the oracle's, both norms, one and two threshold groups; waves
    outside 0..8 are refused."""
    rng = np.random.default_rng(4711)
    w = np.cumsum(rng.standard_normal(2600))
    x = ((w - w.mean()) / w.std()).astype(np.float32)
    try:
        for dist in (1, 2):
            for thr in (np.linspace(0.1, 2.0, 16).astype(np.float32), np.linspace(0.05, 2.5, 21).astype(np.float32)):
                ref = None
                for waves in (7, 4, 1, 0):
                    say("scan waves", dist, thr.size, waves)
                    amd.set_scan_waves(waves)
                    got = amd.histograms(x, 1, 9, thr, dist, which=7)
                    if ref is None:
                        ref = got
                        for k, t in enumerate(thr):
                            assert np.array_equal(got["lam"][k], oracle.hist("lam", x, t, 1, 9, dist)), (dist, t)
                            assert np.array_equal(got["det"][k], oracle.hist("det", x, t, 1, 9, dist)), (dist, t)
                    for key in ("rr", "det", "lam"):
                        assert np.array_equal(ref[key], got[key]), (dist, thr.size, waves, key)
        for bad in (-1, 9):
            with pytest.raises(amd.AccrqaError):
                amd.set_scan_waves(bad)
    finally:
        amd.set_scan_waves(7)


def test_slab_form(amd, oracle):
    """The split form slab by slab (the tiles of one band of rows at a time,
    LAM and DET together, as for series whose tile store does not fit),
    forced with slabs of 256 and 512 rows: the histograms and every measure
    equal the whole-store split form's and the oracle's, both norms, fp32 and
    fp64, both distance forms, several threshold groups, N_c not a multiple of
    the slab height, and DET or LAM asked alone."""
    rng = np.random.default_rng(2025)
    cases = []
    for n, m, tau in ((300, 3, 1), (1000, 16, 1), (2113, 5, 2), (3001, 10, 1)):
        w = np.cumsum(rng.standard_normal(n))
        cases.append((f"walk{n}", (w - w.mean()) / w.std(), m, tau))
    try:
        for name, x64, m, tau in cases:
            for dt in (np.float32, np.float64):
                x = x64.astype(dt)
                thr = np.concatenate([np.linspace(0.05, 2.5, 37), [-1.0]]).astype(dt)
                for dist in (1, 2):
                    for form in (amd.FORM_AUTO, amd.FORM_EXACT):
                        amd.set_form(form)
                        amd.set_slab_rows(0)
                        a = amd.histograms(x, tau, m, thr, dist, which=7)
                        ref = (amd.RR(x, [tau], [m], thr, dist), amd.DET(x, [tau], [m], [2, 3], thr, dist),
                               amd.LAM(x, [tau], [m], [2, 3], thr, dist))
                        for rows in (256, 512):
                            say("slab", name, dt.__name__, dist, form, rows)
                            amd.set_slab_rows(rows)
                            b = amd.histograms(x, tau, m, thr, dist, which=7)
                            for key in ("rr", "det", "lam"):
                                assert np.array_equal(a[key], b[key]), (name, dt, dist, form, rows, key)
                            got = (amd.RR(x, [tau], [m], thr, dist), amd.DET(x, [tau], [m], [2, 3], thr, dist),
                                   amd.LAM(x, [tau], [m], [2, 3], thr, dist))
                            assert same_bits(got[0], ref[0])
                            assert_measures(got[1], ref[1], dt, f"slab DET {name}")
                            assert_measures(got[2], ref[2], dt, f"slab LAM {name}")
                    for k in (3, 20):
                        for kind in ("lam", "det"):
                            want = oracle.hist(kind, x, dt(thr[k]), tau, m, dist)
                            assert np.array_equal(a[kind][k], want), (name, dt, dist, kind, thr[k])
    finally:
        amd.set_slab_rows(0)
        amd.set_form(amd.FORM_AUTO)


def test_mfma_decision_counters(mfma_form, oracle):
    """The MFMA form's diagnostic counters: elements ranked equals the split
    form's column-pass pairs (each band's columns times V_J rows), flagged and
    re-summed counts are ordered, and counting changes no result."""
    amd = mfma_form
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.standard_normal(3000)).astype(np.float32)
    x = ((x - x.mean()) / x.std()).astype(np.float32)
    thr = np.linspace(0.3, 3.0, 16).astype(np.float32)
    tau, m = 1, 16
    nc = x.size - (m - 1) * tau
    base = amd.histograms(x, tau, m, thr, 1, which=6)
    amd.mx_stats(True)
    try:
        got = amd.histograms(x, tau, m, thr, 1, which=6)
        ranked, flagged, exact = amd.mx_stats_read()
    finally:
        amd.mx_stats(False)
    assert np.array_equal(got["det"], base["det"]) and np.array_equal(got["lam"], base["lam"])
    want = sum(min(64, nc - 64 * J) * min(64 * (J + 1), nc) for J in range(-(-nc // 64)))
    assert ranked == want, (ranked, want)
    assert 0 <= exact <= flagged <= ranked, (ranked, flagged, exact)
    assert exact < ranked // 100, (ranked, exact)
    for k, t in enumerate(thr):
        assert np.array_equal(got["lam"][k], oracle.hist("lam", x, np.float32(t), tau, m, 1)), t


def test_rp_golden(amd, golden_rp):
    """accrqa_RP on the GPU against the plots the compiled reference wrote."""
    from conftest import rp_case
    for case in golden_rp["cases"]:
        say("rp golden", case["name"])
        x, tau, emb, thr, dist, ot, want = rp_case(case)
        got = amd.RP(x, tau, emb, thr, dist, ot)
        assert got.dtype == want.dtype and np.array_equal(got, want), case["name"]


def test_rp_against_oracle(amd, oracle):
    """The GPU plot against the oracle's, bit for bit: both norms, fp32 and
    fp64 (char and int output), N_c around the 32-row and 64-column block
    edges, tau = 0, m = 1 and m = 20, NaN and inf samples, ties, and
    thresholds that never recur."""
    rng = np.random.default_rng(41)
    cases = []
    for n, tau, emb in ((31, 1, 1), (64, 1, 2), (97, 2, 3), (130, 0, 4), (700, 3, 20), (1025, 1, 5)):
        cases.append((rng.standard_normal(n), tau, emb))
    w = rng.standard_normal(400)
    w[[7, 300]] = np.nan
    w[100] = np.inf
    cases.append((w, 1, 3))
    cases.append((rng.integers(-3, 4, size=500).astype(np.float64), 2, 2))
    for x64, tau, emb in cases:
        for dt in (np.float32, np.float64):
            x = x64.astype(dt)
            for dist in (1, 2):
                for thr in (0.0, 0.6, 2.0, -1.0):
                    say("rp", x.size, tau, emb, dt.__name__, dist, thr)
                    want, e = oracle.RP(x, tau, emb, dt(thr), dist)
                    assert e == 0
                    got = amd.RP(x, tau, emb, dt(thr), dist)
                    assert np.array_equal(got, want), (x.size, tau, emb, dt, dist, thr)
                    if dt == np.float64 and thr == 0.6:
                        want32, _ = oracle.RP(x, tau, emb, dt(thr), dist, np.int32)
                        assert np.array_equal(amd.RP(x, tau, emb, dt(thr), dist, np.int32), want32)


def test_rp_cxx_device_and_errors(amd, oracle):
    """The reference's mangled accrqa_RP overload (error preset to SUCCESS, as
    the reference requires; a set error makes it do nothing), the device
    entry point on torch buffers, and argument errors."""
    import ctypes as C
    import torch
    L = amd.load()
    rng = np.random.default_rng(43)
    x = rng.standard_normal(600).astype(np.float32)
    want, _ = oracle.RP(x, 2, 3, np.float32(0.7), 1)
    f = getattr(L, "_Z9accrqa_RPPcPfmiif15Accrqa_DistanceP12Accrqa_Error")
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_float, C.c_int, C.c_void_p]
    f.restype = None
    out = np.full(want.shape, 7, np.int8)
    err = C.c_int(0)
    f(out.ctypes.data, x.ctypes.data, x.size, 2, 3, 0.7, 1, C.byref(err))
    assert err.value == 0 and np.array_equal(out, want)
    out[:] = 7
    err = C.c_int(2)
    f(out.ctypes.data, x.ctypes.data, x.size, 2, 3, 0.7, 1, C.byref(err))
    assert err.value == 2 and np.all(out == 7)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    nc = want.shape[0]
    dout = torch.empty(nc * nc, dtype=torch.int8, device=dev)
    amd.rp_device(dx.data_ptr(), x.size, 2, 3, 0.7, 1, dout.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dout.cpu().numpy().reshape(nc, nc), want)
    with pytest.raises(amd.AccrqaError):
        amd.RP(x, 1, 700, 0.5, 1)
    with pytest.raises(amd.AccrqaError):
        amd.RP(x, 1, 3, 0.5, 3)



def test_rr_target_gpu(amd, oracle):
    """RR_target on the GPU (speculative bisection, several thresholds per
    pass) against the oracle's sequential restatement of the reference's
    loop: the same threshold and the same RR bits, both norms, for searches
    that stop by epsilon, by the interval test and by max_iter."""
    from pyoracle import rr_target
    rng = np.random.default_rng(71)
    x = np.cumsum(rng.standard_normal(2500)).astype(np.float32)
    x = (x - x.mean()) / x.std()
    for target, dist, eps, it, lo, hi in ((0.1, 1, 0.01, 20, 0, 10), (0.3, 2, 1e-4, 20, 0, 10),
                                          (0.05, 1, 1e-8, 12, 0, 4), (0.2, 1, 0.01, 20, 6, 10)):
        say("rr_target", target, dist, eps)
        want_t, want_rr, _ = rr_target(oracle, x, 3, 4, target, dist, eps, it, lo, hi)
        got_t, got_rr = amd.RR_target(x, 3, 4, target, dist, eps, it, lo, hi)
        assert got_t == want_t and got_rr.tobytes() == want_rr.tobytes(), (target, dist, eps)


def test_rqa_batch_matches_single_series(amd, oracle):
    """The batched call (persistent batched kernels) against the single-series
    calls (the large-N kernels): RR bit for bit, DET/LAM to the parity bar
    (ENTR, summed in another order, within 1e-6 / 1e-12 relative); series of
    different lengths (one shorter than a 64-column band, one spanning several
    segments of the large-N path), thresholds per series (one never recurs),
    several (tau, emb) -- emb = 11 past the eight column samples the batched
    kernel keeps in registers --, both norms, fp32 and fp64, every measure
    subset; one series also against the oracle directly."""
    rng = np.random.default_rng(73)
    lens = [50, 300, 777, 1500, 2100, 64, 5000]
    taus, embs, lm = [1, 3], [2, 5, 11], [2, 3]
    for dt in (np.float32, np.float64):
        xs = [np.cumsum(rng.standard_normal(n)).astype(dt) for n in lens]
        thr = np.array([[0.1 * s.std(), 0.3 * s.std(), -1.0, 0.7 * s.std()] for s in xs], dtype=dt)
        for dist in (1, 2):
            say("batch", dt.__name__, dist)
            got = amd.rqa_batch(xs, taus, embs, lm, thr, dist)
            for k, s in enumerate(xs):
                rr = amd.RR(s, taus, embs, thr[k], dist)
                det = amd.DET(s, taus, embs, lm, thr[k], dist)
                lam = amd.LAM(s, taus, embs, lm, thr[k], dist)
                assert same_bits(got["RR"][k], rr), (dt, dist, k, "RR")
                assert_measures(got["DET"][k], det, dt, f"batch DET {dt.__name__} {dist} {k}")
                assert_measures(got["LAM"][k], lam, dt, f"batch LAM {dt.__name__} {dist} {k}")
            only = amd.rqa_batch(xs, taus, embs, lm, thr, dist, want=("RR",))
            assert same_bits(only["RR"], got["RR"])
            only = amd.rqa_batch(xs, taus, embs, lm, thr, dist, want=("LAM",))
            assert same_bits(only["LAM"], got["LAM"])
        rr_o, e = oracle.RR(xs[2], taus, embs, thr[2], 1)
        assert e == 0
        assert same_bits(amd.rqa_batch(xs, taus, embs, lm, thr, 1, want=("RR",))["RR"][2].reshape(-1), rr_o)
    with pytest.raises(amd.AccrqaError):
        amd.rqa_batch([np.zeros(10, np.float32), np.zeros(3, np.float32)], [1], [5], [2], [0.1], 1)


def test_rp_wide_form(amd, oracle):
    """The exact-form recurrence plot with four adjacent columns per lane
    (k_rp4: blocks of 16 or 8 rows x 256 columns, one 4-element store per lane
    and row), taken when N_c % 4 == 0 and the plot is aligned: N_c below,
    at and past the block sizes (partial last blocks), against the oracle and
    against one column per lane (accrqa_amd_set_rp_wide(0)), both norms,
    fp32 and fp64, char and int, tau = 0; a plot at an odd device address
    takes the one-column form.  The symmetric form is off here
    (accrqa_amd_set_rp_sym(0)); test_rp_exact_symmetric covers it."""
    import torch
    rng = np.random.default_rng(83)
    amd.set_form(amd.FORM_EXACT)
    amd.set_rp_sym(0)
    try:
        for nc, tau, emb in ((4, 1, 2), (32, 1, 1), (100, 2, 2), (256, 1, 4), (260, 0, 3), (1028, 1, 5),
                             (1040, 3, 9)):
            n = nc + (emb - 1) * tau
            w = np.cumsum(rng.standard_normal(n))
            w = (w - w.mean()) / w.std()
            for dt in (np.float32, np.float64):
                x = w.astype(dt)
                for dist in (1, 2):
                    thr = dt(0.3 * np.sqrt(emb) if dist == 1 else 0.3)
                    want, e = oracle.RP(x, tau, emb, thr, dist)
                    assert e == 0
                    for wide in (1, 0):
                        amd.set_rp_wide(wide)
                        say("rp wide", nc, tau, emb, dt.__name__, dist, wide)
                        assert np.array_equal(amd.RP(x, tau, emb, thr, dist), want), (nc, dt, dist, wide)
                    amd.set_rp_wide(1)
                    if dt == np.float64:
                        want32, _ = oracle.RP(x, tau, emb, thr, dist, np.int32)
                        assert np.array_equal(amd.RP(x, tau, emb, thr, dist, np.int32), want32), (nc, dist)
        # the device entry point into a plot at an odd address
        x = np.cumsum(rng.standard_normal(96 + 2)).astype(np.float32)
        want, _ = oracle.RP(x, 1, 3, np.float32(0.9), 1)
        dx = torch.from_numpy(x).cuda()
        buf = torch.full((96 * 96 + 1,), 7, dtype=torch.int8, device="cuda")
        amd.rp_device(dx.data_ptr(), x.size, 1, 3, 0.9, 1, buf.data_ptr() + 1)
        torch.cuda.synchronize()
        got = buf.cpu().numpy()
        assert got[0] == 7 and np.array_equal(got[1:].reshape(96, 96), want)
    finally:
        amd.set_rp_wide(1)
        amd.set_rp_sym(1)
        amd.set_form(amd.FORM_AUTO)


def test_rp_exact_symmetric(amd, oracle):
    """The exact-form recurrence plot of bytes from the upper triangle
    (k_rp_sym: 64 x 64 squares, each stored as is and mirrored through LDS),
    taken when N_c % 16 == 0 and the plot is 16-byte aligned: N_c of one
    square, partial squares and bands, whole ones, tau = 0, both norms, fp32
    and fp64, against the oracle and against the four-column and one-column
    forms (accrqa_amd_set_rp_sym(0)); a device plot 16 bytes past an aligned
    address takes it, one byte past does not."""
    import torch
    rng = np.random.default_rng(101)
    amd.set_form(amd.FORM_EXACT)
    try:
        for nc, tau, emb in ((16, 1, 3), (48, 2, 2), (64, 1, 5), (128, 0, 3), (208, 1, 4), (576, 3, 9),
                             (1040, 1, 3)):
            n = nc + (emb - 1) * tau
            w = np.cumsum(rng.standard_normal(n))
            w = (w - w.mean()) / w.std()
            for dt in (np.float32, np.float64):
                x = w.astype(dt)
                for dist in (1, 2):
                    thr = dt(0.3 * np.sqrt(emb) if dist == 1 else 0.3)
                    want, e = oracle.RP(x, tau, emb, thr, dist)
                    assert e == 0
                    for sym, wide in ((1, 1), (0, 1), (0, 0)):
                        amd.set_rp_sym(sym)
                        amd.set_rp_wide(wide)
                        say("rp exact sym", nc, tau, emb, dt.__name__, dist, sym, wide)
                        assert np.array_equal(amd.RP(x, tau, emb, thr, dist), want), (nc, dt, dist, sym, wide)
                    amd.set_rp_sym(1)
                    amd.set_rp_wide(1)
        x = np.cumsum(rng.standard_normal(144 + 2)).astype(np.float32)
        want, _ = oracle.RP(x, 1, 3, np.float32(0.9), 1)
        dx = torch.from_numpy(x).cuda()
        for off in (16, 1):
            say("rp exact sym device offset", off)
            buf = torch.full((144 * 144 + 32,), 7, dtype=torch.int8, device="cuda")
            amd.rp_device(dx.data_ptr(), x.size, 1, 3, 0.9, 1, buf.data_ptr() + off)
            torch.cuda.synchronize()
            got = buf.cpu().numpy()
            assert (got[:off] == 7).all() and (got[off + 144 * 144:] == 7).all()
            assert np.array_equal(got[off:off + 144 * 144].reshape(144, 144), want), off
    finally:
        amd.set_rp_sym(1)
        amd.set_rp_wide(1)
        amd.set_form(amd.FORM_AUTO)


def test_rp_mfma_form(amd, oracle):
    """The recurrence plot through the MFMA form (hi/lo Gram, one-bound
    table, exact recheck) against the oracle, bit for bit, also with the form
    forced for m < 8 and off: walks, ties on attainable distances, a spike
    10^6 times the rest, a NaN sample (every pair re-decided exactly), tau = 0,
    m = 33, N_c at the block edges."""
    rng = np.random.default_rng(79)
    w = np.cumsum(rng.standard_normal(700)).astype(np.float32)
    w = (w - w.mean()) / w.std()
    spike = (1e-3 * rng.standard_normal(500)).astype(np.float32)
    spike[222] = 1e3
    nan = rng.standard_normal(400).astype(np.float32)
    nan[99] = np.nan
    ties = rng.integers(-3, 4, size=450).astype(np.float32)
    cases = [(w, 1, 16, 1.5), (w[:96 + 15], 1, 16, 2.0), (w[:161], 2, 9, 0.8), (ties, 1, 8, float(np.sqrt(np.float32(6)))),
             (spike, 1, 9, 4e-3), (nan, 1, 10, 2.5), (w, 0, 12, 0.3), (w, 2, 33, 4.0), (w, 1, 3, 0.4)]
    try:
        for form in (amd.FORM_MFMA, amd.FORM_AUTO, amd.FORM_EXACT):
            amd.set_form(form)
            for x, tau, emb, thr in cases:
                say("rp mfma", form, x.size, tau, emb, thr)
                want, e = oracle.RP(x, tau, emb, np.float32(thr), 1)
                assert e == 0
                got = amd.RP(x, tau, emb, np.float32(thr), 1)
                assert np.array_equal(got, want), (form, x.size, tau, emb, thr)
    finally:
        amd.set_form(amd.FORM_AUTO)


def test_rp_mfma_symmetric(amd, oracle):
    """The MFMA-form recurrence plot from the upper triangle (k_rp_mx_sym:
    64 x 64 squares, each stored as is and mirrored through LDS), taken when
    N_c % 16 == 0 and the plot is 16-byte aligned: N_c of one square, partial
    squares and bands (16, 48, 80, 144, 208, 1040), whole ones (64, 128, 576),
    tau = 0 and 2, m = 33, a spike 10^6 times the rest and a NaN sample (every
    pair re-decided exactly), against the oracle and against every element
    decided (accrqa_amd_set_rp_sym(0)); device plots at an address 16 bytes
    past an aligned one (the mirrored form) and one byte past (every element)."""
    import torch
    rng = np.random.default_rng(97)
    amd.set_form(amd.FORM_MFMA)
    try:
        cases = []
        for nc, tau, emb in ((16, 1, 16), (48, 1, 8), (64, 1, 16), (80, 2, 9), (128, 0, 12), (144, 1, 16),
                             (208, 2, 33), (576, 1, 16), (1040, 3, 10)):
            n = nc + (emb - 1) * tau
            w = np.cumsum(rng.standard_normal(n)).astype(np.float32)
            w = ((w - w.mean()) / w.std()).astype(np.float32)
            cases.append((w, tau, emb, float(0.35 * np.sqrt(emb))))
        spike = (1e-3 * rng.standard_normal(256 + 8)).astype(np.float32)
        spike[77] = 1e3
        cases.append((spike, 1, 9, 4e-3))
        nan = rng.standard_normal(320 + 9).astype(np.float32)
        nan[150] = np.nan
        cases.append((nan, 1, 10, 2.5))
        for x, tau, emb, thr in cases:
            nc = x.size - (emb - 1) * tau
            assert nc % 16 == 0
            want, e = oracle.RP(x, tau, emb, np.float32(thr), 1)
            assert e == 0
            for sym in (1, 0):
                amd.set_rp_sym(sym)
                say("rp sym", nc, tau, emb, thr, sym)
                got = amd.RP(x, tau, emb, np.float32(thr), 1)
                assert np.array_equal(got, want), (nc, tau, emb, thr, sym)
            amd.set_rp_sym(1)
        # device plots 16 bytes and one byte past an aligned address
        x = cases[5][0]
        nc = 144
        want, _ = oracle.RP(x, 1, 16, np.float32(cases[5][3]), 1)
        dx = torch.from_numpy(x).cuda()
        for off in (16, 1):
            say("rp sym device offset", off)
            buf = torch.full((nc * nc + 32,), 7, dtype=torch.int8, device="cuda")
            amd.rp_device(dx.data_ptr(), x.size, 1, 16, cases[5][3], 1, buf.data_ptr() + off)
            torch.cuda.synchronize()
            got = buf.cpu().numpy()
            assert (got[:off] == 7).all() and (got[off + nc * nc:] == 7).all()
            assert np.array_equal(got[off:off + nc * nc].reshape(nc, nc), want), off
    finally:
        amd.set_rp_sym(1)
        amd.set_form(amd.FORM_AUTO)


def test_compat_module_on_gpu(amd, oracle):
    """accrqa_amd.compat (the reference's Python interface) on the GPU: its
    cubes are the oracle's flat outputs in the reference's shapes, bit for
    bit, and RP / RR_target go through."""
    import accrqa_amd.compat as accrqa
    rng = np.random.default_rng(89)
    x = np.cumsum(rng.standard_normal(900)).astype(np.float32)
    x = ((x - x.mean()) / x.std()).astype(np.float32)
    for name in ("euclidean", "maximal"):
        d = accrqa.accrqaDistance(name)
        did = d.get_distance_id()
        rr = accrqa.RR(x, [1, 3], [2, 4], [0.2, 0.5], d, tidy_data=False)
        rr_o, _ = oracle.RR(x, [1, 3], [2, 4], [0.2, 0.5], did)
        assert same_bits(rr.reshape(-1), rr_o.reshape(-1))
        for kind in ("DET", "LAM"):
            got = getattr(accrqa, kind)(x, [1, 3], [2, 4], [2, 3], [0.2, 0.5], d, tidy_data=False)
            exp, _ = getattr(oracle, kind)(x, [1, 3], [2, 4], [2, 3], [0.2, 0.5], did)
            exp = np.asarray(exp).reshape(-1, 5)
            g = got.reshape(-1, 5)
            for k in (0, 1, 2, 4):
                assert same_bits(g[:, k].copy(), exp[:, k].copy()), (name, kind, k)
        tidy = accrqa.DET(x, [1], [2], [2], [0.3], d)
        assert len(tidy) == 1 and tidy["Lmin"].iloc[0] == 2
    rp = accrqa.RP(x[:300], 1, 3, 0.4, accrqa.accrqaDistance("euclidean"))
    rp_o, _ = oracle.RP(x[:300], 1, 3, np.float32(0.4), 1)
    assert np.array_equal(rp, rp_o)
    from pyoracle import rr_target
    t, r = accrqa.RR_target(x, 1, 3, 0.1, accrqa.accrqaDistance("euclidean"))
    want_t, want_rr, _ = rr_target(oracle, x, 1, 3, 0.1, 1, 0.01, 20, 0, 10)
    assert t == want_t and np.float32(r).tobytes() == want_rr.tobytes(), (t, r, want_t, want_rr)


def test_fp64_mismatch_diagnostic(amd, oracle):
    """The fp32-vs-true-fp64 decision counter against the oracle's two plots:
    the reference's fp32 decisions and its fp64 decisions on the same fp32
    samples and thresholds, differing pairs counted over the full matrix.
    Thirds of integers are not fp32 values, so many distances land within
    rounding of the thresholds (both norms, ties and near ties)."""
    import torch
    rng = np.random.default_rng(5)
    x = (rng.integers(-9, 10, size=900) / 3.0).astype(np.float32)
    m = 3
    dx = torch.from_numpy(x).cuda()
    for dist in (1, 2):
        thr = [np.float32(np.sqrt(k) / 3.0) if dist == 1 else np.float32(k / 3.0) for k in (1, 2, 3, 5, 8, 13)]
        got = amd.fp64_mismatches(dx.data_ptr(), x.size, 1, m, thr, dist)
        want = []
        for t in thr:
            a, _ = oracle.RP(x, 1, m, t, dist)
            b, _ = oracle.RP(x.astype(np.float64), 1, m, np.float64(t), dist)
            want.append(int((a != b).sum()))
        say("fp64 diag", dist, list(got), want)
        assert [int(v) for v in got] == want, (dist, got, want)
        assert sum(want) > 0
