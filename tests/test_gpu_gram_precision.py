"""The Gram precisions of the split MFMA form's rank pass (accrqa_amd_set_gram):
one FP16 half, one BF16 half or one FP8 (E4M3) byte per sample instead of the
FP16 hi/lo pair.  Each has its own rigorous error bound (engine mx_params), so
the table, the second stage and the exact re-decision must still reproduce
every decision of the reference: histograms and measures against the oracle,
bit for bit (ENTR to the parity bar), on walks, ties, a spike, tiny and huge
series.  The probe checks what those bounds take of the matrix instructions:
the operand rounding within the stated eps and eta, and the result of c + qa qb
within 32u (|c| + sum |p~|) for the 16-bit formats and 2^-9 (|c| + sum |p~|)
for FP8, p~ counting subnormal operands at the smallest normal (exact on
small integers)."""
import numpy as np
import pytest

from conftest import check_case, say

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
G16 = 32 * U / (1 - 32 * U)


@pytest.fixture(scope="module")
def amd():
    import accrqa_amd
    accrqa_amd.load()
    return accrqa_amd


# Smallest normal of each operand format: a subnormal operand's product is
# aligned by that exponent, so the accumulation bound counts it at that size.
TINY = {1: 2.0 ** -14, 2: 2.0 ** -126, 3: 2.0 ** -6}


def acc_bound(gram, A, B, C):
    """The accumulation bound the Gram's error analysis takes per instruction:
    16-bit operands 32u (|c| + sum |p~|), FP8 2^-9 (|c| + sum |p~|), with p~ the
    products of the operands raised to the format's smallest normal."""
    t = TINY[gram]
    nA = np.where(A == 0, 0.0, np.maximum(np.abs(A), t))
    nB = np.where(B == 0, 0.0, np.maximum(np.abs(B), t))
    k = 2.0 ** -9 if gram == 3 else G16
    return k * (np.abs(C) + nA @ nB) * (1 + 2.0 ** -40)


def _bf16_rne(a):
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("gram", [1, 2, 3])
def test_gram_probe_properties(amd, gram):
    rng = np.random.default_rng(11 + gram)
    # (eps, eta) of the bound taken for each precision
    eps, eta = {1: (2.0 ** -11, 2.0 ** -25), 2: (2.0 ** -8, 2.0 ** -126), 3: (2.0 ** -3, 2.0 ** -6)}[gram]
    top = 2.0 ** 8 if gram == 3 else 2.0 ** 14   # scaled samples stay below 2^amax
    worst = 0.0
    for s in range(6):
        if s == 0:
            a = rng.integers(-8, 9, (32, 16)).astype(np.float32)
            b = rng.integers(-8, 9, (16, 32)).astype(np.float32)
        else:
            ea = rng.integers(-12 if gram == 3 else -30, 7 if gram == 3 else 13, (32, 16))
            eb = rng.integers(-12 if gram == 3 else -30, 7 if gram == 3 else 13, (16, 32))
            a = (rng.choice([-1.0, 1.0], (32, 16)) * (1 + rng.random((32, 16))) * 2.0 ** ea).astype(np.float32)
            b = (rng.choice([-1.0, 1.0], (16, 32)) * (1 + rng.random((16, 32))) * 2.0 ** eb).astype(np.float32)
        assert np.all(np.abs(a) < top) and np.all(np.abs(b) < top)
        c = (rng.standard_normal((32, 32)) * (0 if s == 0 else 2.0 ** rng.integers(-10, 10, (32, 32)))).astype(np.float32)
        d, qa, qb = amd.gram_probe(gram, a, b, c)
        for x, q in ((a, qa), (b, qb)):
            err = np.abs(q.astype(np.float64) - x.astype(np.float64))
            assert np.all(err <= eps * np.abs(x.astype(np.float64)) + eta), (gram, s, err.max())
            big = np.abs(x) >= 2.0 ** -6
            if np.any(big):
                worst = max(worst, float((err[big] / np.abs(x[big].astype(np.float64))).max()))
        if gram == 2:
            assert np.array_equal(qa, _bf16_rne(a)) and np.array_equal(qb, _bf16_rne(b)), s
        if gram == 1:
            assert np.array_equal(qa, a.astype(np.float16).astype(np.float32)), s
        A, B, Cc = qa.astype(np.float64), qb.astype(np.float64), c.astype(np.float64)
        want = Cc + A @ B
        if s == 0 and gram != 3:
            assert np.array_equal(d.astype(np.float64), want), (gram, "integers")
        bound = acc_bound(gram, A, B, Cc)
        err = np.abs(d.astype(np.float64) - want)
        bad = np.argwhere(err > bound)
        for i, j in bad[:8]:
            p = A[i, :] * B[:, j]
            say("gram", gram, "case", s, "element", (i, j), "d", float(d[i, j]), "want", want[i, j],
                "err", err[i, j], "bound", bound[i, j], "c", Cc[i, j], "sum|p|", np.abs(p).sum(),
                "max|p|", np.abs(p).max(), "min|p|>0", np.abs(p[p != 0]).min() if np.any(p != 0) else 0.0,
                "products", list(p))
        assert bad.size == 0, (gram, s, len(bad))
    say("gram", gram, "largest relative rounding error of normal operands", worst)


def test_gram_precisions_bit_exact(amd, oracle):
    rng = np.random.default_rng(91)
    cases = []
    x = np.cumsum(rng.standard_normal(2600)).astype(np.float32)
    cases.append(("walk", ((x - x.mean()) / x.std()).astype(np.float32), [1, 2], [16, 9], [2, 3],
                  list(np.linspace(0.2, 3.0, 16).astype(np.float32))))
    cases.append(("ties", rng.integers(-4, 5, size=1100).astype(np.float32), [1], [8, 12], [2],
                  [np.float32(np.sqrt(np.float32(k))) for k in (1, 4, 9, 17, 40, 61)]))
    xs = (1e-3 * rng.standard_normal(1500)).astype(np.float32)
    xs[700] = 1e3
    cases.append(("spike", xs, [1], [16], [2], list(np.linspace(2e-4, 8e-3, 11).astype(np.float32))))
    cases.append(("tiny", (1e-18 * rng.standard_normal(1000)).astype(np.float32), [1], [16], [2],
                  list(np.linspace(1e-19, 6e-18, 9).astype(np.float32))))
    cases.append(("huge", (1e12 * rng.standard_normal(1000)).astype(np.float32), [2], [20], [2],
                  list(np.linspace(1e11, 6e12, 9).astype(np.float32))))
    amd.set_form(amd.FORM_MFMA)
    rechecks = {}
    try:
        for gram in (1, 2, 3):
            amd.set_gram(gram)
            for name, x, taus, embs, lm, thrs in cases:
                say("gram", gram, name)
                check_case(amd, oracle, x, taus, embs, lm, thrs, 1, f"gram {gram} {name}")
                got = amd.histograms(x, taus[0], embs[0], np.asarray(thrs, np.float32), 1, which=7)
                for k, t in enumerate(thrs):
                    t = np.float32(t)
                    assert np.array_equal(got["lam"][k], oracle.hist("lam", x, t, taus[0], embs[0], 1)), (gram, name, t)
                    assert np.array_equal(got["det"][k], oracle.hist("det", x, t, taus[0], embs[0], 1)), (gram, name, t)
            # the share of pairs re-decided in exact order grows as the precision drops
            name, x, taus, embs, lm, thrs = cases[0]
            amd.mx_stats(True)
            amd.histograms(x, 1, 16, np.asarray(thrs, np.float32), 1, which=7)
            ranked, flagged, exact = amd.mx_stats_read()
            amd.mx_stats(False)
            rechecks[gram] = exact / ranked
        amd.set_gram(0)
        amd.mx_stats(True)
        amd.histograms(cases[0][1], 1, 16, np.asarray(cases[0][5], np.float32), 1, which=7)
        ranked, flagged, exact = amd.mx_stats_read()
        amd.mx_stats(False)
        rechecks[0] = exact / ranked
        say("recheck fractions", rechecks)
        assert rechecks[0] <= rechecks[1] <= rechecks[2] <= rechecks[3]
        # more than 16 thresholds per block: refused while a lower precision is set
        amd.set_gram(1)
        with pytest.raises(amd.AccrqaError):
            amd.LAM(cases[0][1], [1], [16], [2], np.linspace(0.2, 3.0, 20).astype(np.float32), 1)
    finally:
        amd.set_gram(0)
        amd.set_form(amd.FORM_AUTO)
