"""Generate the golden fixtures under tests/golden/ from the REFERENCE CPU path.

Run in the development container, where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

The reference's CPU path is compiled from its own sources by oracle/Makefile
into oracle/_ref/libaccrqa_ref.so; this script calls it (public API plus the
raw-histogram glue in oracle/ref_glue.cpp) on small seeded inputs and writes
inputs and outputs as JSON.  Floats are stored as hex strings (float.hex) so
that comparisons are bit-exact; histograms are stored sparsely as
{length: count}.  Only the fixtures are committed; the reference never
travels.

Two data files the reference's own tests and example hold are copied verbatim
as inputs: tests/testthat/testdata/input.txt (the 1,122-sample series of the
CRQA known-answer tests) and application/test_data.txt (1,000 samples).
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from pyoracle import Reference, nc_of  # noqa: E402

REF_TREE = "/root/reference"


def hx(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).reshape(-1)]


def sparse(h):
    nz = np.nonzero(h)[0]
    return {str(int(k)): int(h[k]) for k in nz}


def case_record(R, name, x, taus, embs, lmins, thrs, dist):
    dt = x.dtype
    rec = {
        "name": name,
        "dtype": "f64" if dt == np.float64 else "f32",
        "x": hx(x),
        "tau": [int(t) for t in taus],
        "emb": [int(m) for m in embs],
        "lmin": [int(v) for v in lmins],
        "thr": hx(np.asarray(thrs, dtype=dt)),
        "dist": int(dist),
    }
    rr, e1 = R.RR(x, taus, embs, thrs, dist)
    det, e2 = R.DET(x, taus, embs, lmins, thrs, dist)
    lam, e3 = R.LAM(x, taus, embs, lmins, thrs, dist)
    assert e1 == e2 == e3 == 0, (name, e1, e2, e3)
    rec["RR"] = hx(rr)
    rec["DET"] = hx(det)
    rec["LAM"] = hx(lam)
    hists = []
    for tau in taus:
        for emb in embs:
            for t in np.asarray(thrs, dtype=dt):
                hd = R.hist("det", x, t, tau, emb, dist)
                hv = R.hist("lam", x, t, tau, emb, dist)
                hists.append({"tau": int(tau), "emb": int(emb), "thr": float(t).hex(),
                              "nc": nc_of(x.size, tau, emb), "hD": sparse(hd), "hV": sparse(hv)})
    rec["hist"] = hists
    return rec


def logistic(n, seed, dt):
    rng = np.random.default_rng(seed)
    v = rng.uniform(0.1, 0.9)
    out = np.empty(n)
    for k in range(n):
        out[k] = v
        v = 4.0 * v * (1.0 - v)
    return out.astype(dt)


def transcribe_crqa_kats():
    """Copy the CRQA known-answer numbers out of the reference's C++ test
    (tests/test_RQA.cpp:1179-1359) into crqa_kat.json.  Only the numbers are
    kept (tau, emb, l_min/v_min, threshold, expected DET/L/ENTR/LAM/TT/RR)."""
    import re
    src = open(os.path.join(REF_TREE, "tests/test_RQA.cpp")).read()

    def grab(fn_name):
        start = src.index("void " + fn_name + "(")
        end = src.find("\nvoid ", start + 10)
        if end < 0:
            end = src.find("\nbool ", start + 10)
        out, norm = [], "euclidean"
        for line in src[start:end].splitlines():
            if "DST_MAXIMAL" in line:
                norm = "maximal"
            if "tau" in line and "thr" in line and "CRQA_" in line and "Run_" not in line and "double" not in line:
                rec = {"norm": norm}
                for k, v in re.findall(r"(\w+)\s*=\s*([-0-9.e]+)", line):
                    rec[k] = float(v)
                out.append(rec)
        return out

    kat = {"source": "CRQA known-answer values held by the reference's C++ tests (tests/test_RQA.cpp:1179-1359), "
                     "series tests/testthat/testdata/input.txt (1,122 samples, double), l_min = v_min = 2, "
                     "tolerance 1e-4 absolute",
           "DET": grab("test_against_CRQA_DET"), "LAM": grab("test_against_CRQA_LAM"),
           "RR": grab("test_against_CRQA_RR")}
    with open(os.path.join(HERE, "crqa_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


def rp_record(R, name, x, tau, emb, thr, dist, out_dtype=np.int8):
    """One accrqa_RP call: the plot's bits packed row-major (np.packbits of
    plot != 0, hex) and its sum."""
    dt = x.dtype
    thr = dt.type(thr)
    rp, e = R.RP(x, tau, emb, thr, dist, out_dtype)
    assert e == 0, (name, e)
    assert set(np.unique(rp)) <= {0, 1}, name
    return {"name": name, "dtype": "f64" if dt == np.float64 else "f32", "x": hx(x), "tau": int(tau),
            "emb": int(emb), "thr": float(thr).hex(), "dist": int(dist),
            "out": "i32" if np.dtype(out_dtype) == np.int32 else "i8", "nc": int(rp.shape[0]),
            "sum": int(rp.sum(dtype=np.int64)), "bits": np.packbits(rp.reshape(-1) != 0).tobytes().hex()}


def rp_cases(R, app):
    """Recurrence plots of accrqa_RP (its three overloads), both norms."""
    rng = np.random.default_rng(20261203)
    cases = []
    a = app[:260]
    for dist in (1, 2):
        cases.append(rp_record(R, f"app_f32_d{dist}", a.astype(np.float32), 1, 3, 0.3, dist))
        cases.append(rp_record(R, f"app_f64_d{dist}", a.astype(np.float64), 2, 4, 0.4, dist))
        cases.append(rp_record(R, f"app_f64_i32_d{dist}", a.astype(np.float64), 1, 2, 0.25, dist, np.int32))
    x = rng.standard_normal(220)
    x[31] = np.nan
    x[150] = np.inf
    for dt in (np.float32, np.float64):
        for thr in (-1.0, 0.0, 0.8, np.inf):
            cases.append(rp_record(R, f"special_{np.dtype(dt).name}_{thr}", x.astype(dt), 1, 2, thr, 1))
    ties = rng.integers(-3, 4, size=240).astype(np.float32)
    cases.append(rp_record(R, "ties_f32", ties, 1, 3, float(np.sqrt(np.float32(5.0))), 1))
    cases.append(rp_record(R, "ties_f32_tau0", ties, 0, 3, 2.0, 1))
    cases.append(rp_record(R, "walk_f32_m16", np.cumsum(rng.standard_normal(300)).astype(np.float32), 1, 16,
                           4.0, 1))
    out = os.path.join(HERE, "golden_rp.json")
    with open(out, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "source": "reference accrqa_RP (src/AccRQA_functions.cpp:996-1060) compiled by oracle/Makefile",
                   "cases": cases}, f, separators=(",", ":"))
    print(f"wrote {len(cases)} RP cases to {out} ({os.path.getsize(out) / 1e3:.0f} kB)")


def main():
    transcribe_crqa_kats()
    if not Reference.available():
        sys.exit("oracle/_ref/libaccrqa_ref.so missing: run `make -C oracle ref` first")
    R = Reference()

    # Data files held by the reference, copied verbatim.
    shutil.copyfile(os.path.join(REF_TREE, "tests/testthat/testdata/input.txt"),
                    os.path.join(HERE, "crqa_input.txt"))
    shutil.copyfile(os.path.join(REF_TREE, "application/test_data.txt"),
                    os.path.join(HERE, "app_test_data.txt"))
    crqa = np.array([float(v) for v in open(os.path.join(HERE, "crqa_input.txt")).read().split()])
    app = np.array([float(v) for v in open(os.path.join(HERE, "app_test_data.txt")).read().split()])

    rng = np.random.default_rng(20261202)
    cases = []
    # 1. CRQA series (double, as the reference's KATs use it) and a float copy.
    for dt in (np.float64, np.float32):
        for dist in (1, 2):
            cases.append(case_record(R, f"crqa_{np.dtype(dt).name}_d{dist}", crqa.astype(dt), [6, 10], [3, 4],
                                     [2, 5], [1.2, 1.8], dist))
    # 2. Application data with a spread of thresholds.
    for dist in (1, 2):
        cases.append(case_record(R, f"app_f32_d{dist}", app.astype(np.float32), [1, 3], [2, 5], [2, 3, 7],
                                 [0.0, 0.05, 0.1, 0.2, 0.35], dist))
    # 3. Logistic map (the CPU-runnable configuration's generator), small N.
    for dt in (np.float32, np.float64):
        cases.append(case_record(R, f"logistic_{np.dtype(dt).name}", logistic(1500, 7, dt), [1], [3], [2],
                                 [0.05, 0.1, 0.15, 0.2, 0.25], 1))
    # 4. Random normal data, assorted shapes, both norms and dtypes.
    for k in range(8):
        dt = np.float32 if k % 2 == 0 else np.float64
        n = int(rng.integers(200, 900))
        x = rng.standard_normal(n).astype(dt)
        taus = [int(rng.integers(1, 5))]
        embs = [int(rng.integers(1, 9))]
        thrs = sorted(float(v) for v in rng.uniform(0.2, 3.0, size=4))
        cases.append(case_record(R, f"normal_{k}", x, taus, embs, [1, 2, 4], thrs, 1 + (k % 3 == 0)))
    # 5. Ties: integer-valued data, thresholds exactly at attainable distances
    #    (squared distances are small integers, so sqrt(s) == eps is exact).
    for dt in (np.float32, np.float64):
        x = rng.integers(-3, 4, size=600).astype(dt)
        for dist in (1, 2):
            cases.append(case_record(R, f"ties_{np.dtype(dt).name}_d{dist}", x, [1, 2], [2, 3], [2, 3],
                                     [0.0, 1.0, 2.0, float(np.sqrt(dt(5.0))), 3.0], dist))
    # 6. Special thresholds: negative, -0, +inf, tiny, and data with a NaN and an inf.
    for dt in (np.float32, np.float64):
        x = rng.standard_normal(400).astype(dt)
        x[37] = np.nan
        x[250] = np.inf
        for dist in (1, 2):
            cases.append(case_record(R, f"special_{np.dtype(dt).name}_d{dist}", x, [1], [2], [2],
                                     [-1.0, -0.0, 0.0, 1e-30, 0.5, np.inf], dist))
    # 7. Constant series (closed forms), tiny and longer.
    for dt in (np.float32, np.float64):
        cases.append(case_record(R, f"const_{np.dtype(dt).name}", np.ones(10, dtype=dt), [1, 2], [1],
                                 [2, 3, 4], [0.5], 1))
    # 8. Slowly varying signal with long lines (sine + small noise), float32.
    t = np.arange(2000)
    x = (np.sin(2 * np.pi * t / 97.0) + 0.01 * rng.standard_normal(2000)).astype(np.float32)
    cases.append(case_record(R, "sine_f32", x, [1, 5], [3], [2, 10], [0.05, 0.2, 0.6], 1))
    cases.append(case_record(R, "sine_f32_max", x, [1], [3, 6], [2], [0.05, 0.2, 0.6], 2))

    out = os.path.join(HERE, "golden_reference.json")
    with open(out, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "source": "reference CPU path (src/AccRQA_functions.cpp, src/AccRQA_CPU_function.cpp) "
                             "compiled by oracle/Makefile",
                   "cases": cases}, f, separators=(",", ":"))
    print(f"wrote {len(cases)} cases to {out} ({os.path.getsize(out) / 1e3:.0f} kB)")
    rp_cases(R, app)


if __name__ == "__main__":
    main()
