#!/usr/bin/env python3
"""Headline measurement: RR + DET + LAM logical pair-evaluations per second on one MI355X.

One step = one full pass of the path over one synthetic series: RR, DET
(DET, L, Lmax, ENTR, RR) and LAM (LAM, TT, TTmax, ENTR, RR) for every
threshold of the configuration, with the series already resident in HBM
(accrqa_amd_rqa_device_*).  A logical pair-evaluation is one recurrence
decision R(i, j) for one threshold over the full N_c x N_c matrix, so
value = N_c^2 * T / (seconds per step).

Default workload: AR(1) phi = 0.95 plus a 0.3-amplitude sine of period 1,000,
N = 262,144, z-normalised, fp32, tau = 1, m = 16, 16 thresholds
linspace(0.1, 1.6) * sqrt(m), Euclidean norm, l_min = 2.

Timing: W warm-up steps, then K steps between two stream synchronisations
with nothing else in the loop (ms_per_step, value); then the same K steps
again with the library's per-kernel HIP events on (an event pair around each
launch on its own stream, read back after the loop), which give each
kernel's average launch duration.  In the timed steps the split form runs its
DET tile scan on a second stream beside the LAM column and row scans
(accrqa_amd_set_overlap; --no-overlap runs them one after the other); the
per-kernel pass always runs them one after the other, so each kernel's
duration is its own.

The JSON line also carries
  roofline     -- the dominant kernel of the reported form (the largest one
                  on the step's critical path: with the overlap, k_det_scan
                  runs beside the longer LAM scans and is off it), its
                  algorithmic work per launch over its
                  average launch duration against the published peak of the
                  unit that work runs on: k_lam_mx, the MFMA form's Gram and
                  ranks, is priced at 2m FP16 FLOP per distinct pair
                  (N_c(N_c+1)/2 per launch) against 2.5 PFLOP/s; the exact
                  form's k_lam at 3m FP32 operations per distinct pair against
                  the VALU; the tile scans at the bytes of the rank tiles they
                  read against 8 TB/s.  Every kernel's figures are under
                  forms.*.rooflines;
  cpu_baseline -- the reference CPU path (oracle/_ref, OpenMP over this
                  host's cores, bound spread) on a bounded prefix of the same
                  series: one warm-up, the median of three timed runs.
"""
import os

# The CPU baseline (OpenMP) is bound spread over the host's cores; set before
# any library that starts an OpenMP runtime is loaded.
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

import argparse  # noqa: E402
import json  # noqa: E402
import statistics  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X FP32 VALU: two wave64 instructions per CU per clock (128 lanes x
# 256 CUs x 2.4 GHz = 78.6 T lane-ops/s), the rate an add/mul probe with
# VGPR operands measures (120.8 lane-ops per CU per clock); the published
# 157.3 TFLOP/s is that rate with an FMA counted as 2.  The decision
# arithmetic is exact-order sub/mul/add/compare (no FMA, no packed FP32), so
# one op per lane-slot is the ceiling.  An instruction with an SGPR operand
# issues at half that rate (the same probe: 62 lane-ops per CU per clock).
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
PUBLISHED_FP32_TFLOPS = 157.3
PUBLISHED_FP16_TFLOPS = 2500.0   # dense FP16 MFMA


def workload(name):
    from accrqa_amd import workloads as W
    if name == "ar1_m16":
        c = W.config3(m=16)
        c["desc"] = ("AR(1) phi=0.95 + 0.3 sin(2 pi t/1000), N=262144, z-normalised, fp32, tau=1, m=16, "
                     "16 thresholds linspace(0.1,1.6)*sqrt(16), Euclidean, l_min=2")
    elif name == "lorenz":
        c = W.config2()
        c["norms"] = [1]
        c["desc"] = ("Lorenz x (RK4 dt=0.01), N=131072, z-normalised, fp32, tau=8, m=3, 20 thresholds "
                     "log-spaced 0.02..0.5, Euclidean, l_min=2")
    elif name == "logistic":
        c = W.config1()
        c["desc"] = "logistic map, N=10000, fp32, tau=1, m=3, 5 thresholds 0.05..0.25, Euclidean, l_min=2"
    else:
        raise SystemExit(f"unknown workload {name}")
    return c


def tile_store_bytes(nc):
    """Bytes of the upper-triangle rank tiles (32 x 64 pairs, one byte each)."""
    nrb, nb = -(-nc // 32), -(-nc // 64)
    return (nrb * nb - (nrb // 2) * ((nrb - 1) // 2)) * 2048


# HBM bytes per launch of the kernels on the default workload (value, source):
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over one step of the MFMA
# form, scaled by the same counters' readings for a 1 GiB device copy
# (FETCH_SIZE reads 0.5 GiB for it: x2.0; WRITE_SIZE 1 GiB: x1.0).
_TRAFFIC_SRC = "profiles/ar1_m16_traffic.md"
MEASURED_TRAFFIC = {
    ("ar1_m16", "k_lam_mx"): (2.0 * 0.9967e9 + 34.373e9, _TRAFFIC_SRC),
    ("ar1_m16", "k_lam_vscan"): (2.0 * 17.305e9 + 5.311e9, _TRAFFIC_SRC),
    ("ar1_m16", "k_lam_hscan"): (2.0 * 17.287e9 + 5.291e9, _TRAFFIC_SRC),
    ("ar1_m16", "k_det_scan"): (2.0 * 25.343e9 + 4.015e9, _TRAFFIC_SRC),
}


# VALU issue of the kernels on the default workload (rocprofv3 --pmc
# SQ_INSTS_VALU and GRBM_GUI_ACTIVE over one step): wave64 VALU instructions
# per launch.  An issue-rate probe puts the ceiling of their instruction mix at
# 0.99-1.01 per CU per cycle for shifts, shift-ors and find-first-bit, 1.72 for
# v_bitop3 and 1.87 for FP32 adds.
_ISSUE_SRC = "profiles/ar1_m16_instruction_mix.md"
MEASURED_VALU = {
    ("ar1_m16", "k_lam_vscan"): 7.029e9,
    ("ar1_m16", "k_lam_hscan"): 7.298e9,
    ("ar1_m16", "k_det_scan"): 7.122e9,
    ("ar1_m16", "k_lam_mx"): 6.634e9,
    ("ar1_m16", "k_lam"): 3.921e10,
}
CU_COUNT, CLOCK_GHZ = 256, 2.4


def kernel_rooflines(kern, nc, emb, T, dist, workload, tile_store):
    """Roofline figures of the main kernels from their average launch times.
    Algorithmic work per launch (the distinct pairs are the N_c(N_c+1)/2 of
    the upper triangle and the diagonal; each is decided once per launch):
      k_lam_mx     2m FP16 MFMA FLOP per distinct pair (the Gram x_i . x_j)
                   against the dense FP16 peak; beside it the hi/lo form's
                   3 x 2m issued FLOP and the bytes of rank tiles it writes;
      k_lam        3m (Euclidean) or 2m (maximal) exact-order FP32 VALU
                   operations plus the T threshold compares per distinct pair
                   against the VALU issue rate (VALU_PEAK_TOPS);
      k_lam_vscan, one read of the rank-tile store (the form's intermediate,
      k_lam_hscan, one byte per distinct pair) against 8 TB/s.
      k_det_scan"""
    pairs = nc * (nc + 1) / 2.0
    out = {}
    for k, v in kern.items():
        t = v["ms_total"] / max(v["launches"], 1) / 1e3
        if t <= 0:
            continue
        if k == "k_lam_mx":
            g = pairs * 2 * emb
            out[k] = {"bound": "mfma", "achieved": g / t / 1e12, "peak": PUBLISHED_FP16_TFLOPS, "unit": "TFLOP/s",
                      "work_per_launch": g, "work_unit": "FP16 MFMA FLOP, 2m per distinct pair",
                      "issued_hilo": {"flop_per_launch": 3 * g, "achieved": 3 * g / t / 1e12,
                                      "frac": 3 * g / t / 1e12 / PUBLISHED_FP16_TFLOPS,
                                      "note": "h.h + h.l + l.h products issued (3 x 2m per distinct pair)"}}
            if tile_store:
                w = float(tile_store_bytes(nc))
                out[k]["tile_store_write"] = {"bytes": w, "GB_s": w / t / 1e9, "frac_of_8TB_s": w / t / 8e12}
        elif k == "k_lam":
            w = pairs * ((3 * emb if dist == 1 else 2 * emb) + T)
            w_sum = pairs * (3 * emb if dist == 1 else 2 * emb)
            out[k] = {"bound": "valu", "achieved": w / t / 1e12, "peak": VALU_PEAK_TOPS, "unit": "TFLOP/s",
                      "work_per_launch": w,
                      "work_unit": ("FP32 VALU operations per distinct pair: the exact-order sum (%s) and "
                                    "the T threshold compares" % ("3m" if dist == 1 else "2m")),
                      "sum_only": {"work_per_launch": w_sum, "achieved": w_sum / t / 1e12,
                                   "frac": w_sum / t / 1e12 / VALU_PEAK_TOPS},
                      "peak_note": "FP32 VALU non-FMA issue rate, VGPR operands (sub/mul/add in order, no FMA)"}
        elif k in ("k_lam_vscan", "k_lam_hscan", "k_det_scan") and tile_store:
            w = float(tile_store_bytes(nc))
            out[k] = {"bound": "hbm", "achieved": w / t / 1e9, "peak": 8000.0, "unit": "GB/s",
                      "work_per_launch": w, "work_unit": "bytes (one read of the rank-tile store)",
                      "note": ("the rank-tile store is this design's intermediate (the path's own traffic is "
                               "under 0.1 GB per step); the scans are bound by VALU issue (valu_issue)")}
        else:
            continue
        r = out[k]
        r["frac"] = r["achieved"] / r["peak"]
        r["avg_ms"] = t * 1e3
        vi = MEASURED_VALU.get((workload, k))
        if vi:
            r["valu_issue"] = {"wave_instructions_per_launch": vi,
                               "per_cu_per_cycle": vi / (CU_COUNT * t * CLOCK_GHZ * 1e9),
                               "source": _ISSUE_SRC}
        r["launches"] = v["launches"]
        tr = MEASURED_TRAFFIC.get((workload, k))
        r["traffic"] = tr[0] if tr else None
        if tr:
            r["traffic_note"] = ("HBM bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the "
                                 "counters' reading for a device copy of known size; " + tr[1])
    return out


_T0 = time.perf_counter()


def progress(what):
    """One line on stderr as each part of the run ends (the JSON line is the
    only stdout)."""
    print(f"bench: {what} done at {time.perf_counter() - _T0:.1f} s", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x, tau, emb, thr, lmin, dist, budget_s):
    """Times the reference CPU path (or the oracle where the reference build
    is absent) on a prefix of the series sized so that one RR+DET+LAM run
    takes about budget_s / 4 seconds: one warm-up run, then the median of
    three timed runs.  Returns the JSON object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, Reference
    O = Oracle()
    if Reference.available():
        impl, kind = Reference(), "reference"
    else:
        impl, kind = O, "port"
    T = len(thr)

    def run(ns):
        xs = np.ascontiguousarray(x[:ns])
        t0 = time.perf_counter()
        impl.RR(xs, [tau], [emb], thr, dist)
        impl.DET(xs, [tau], [emb], lmin, thr, dist)
        impl.LAM(xs, [tau], [emb], lmin, thr, dist)
        return time.perf_counter() - t0

    per_run = budget_s / 4.0
    ns = min(len(x), 1024 + (emb - 1) * tau)
    t = run(ns)
    while t < 0.3 and ns < len(x):
        ns = min(len(x), ns * 2)
        t = run(ns)
    if ns < len(x):
        nc0 = ns - (emb - 1) * tau
        ns = min(len(x), int(nc0 * (per_run / max(t, 1e-3)) ** 0.5) + (emb - 1) * tau)
    run(ns)   # warm-up at the final size
    times = [run(ns) for _ in range(3)]
    t = statistics.median(times)
    nc = ns - (emb - 1) * tau
    return {
        "value": float(nc) * nc * T / t,
        "unit": "pair-evals/s",
        "cores": O.num_threads(),
        "kind": kind,
        "sample": (f"prefix of {ns} samples (N_c={nc}) of the same series, RR+DET+LAM via accrqa_RR/DET/LAM "
                   f"PLT_CPU, {T} thresholds; median of 3 runs after a warm-up: "
                   + ", ".join(f"{v:.2f}" for v in times) + " s"),
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ar1_m16", choices=["ar1_m16", "lorenz", "logistic"])
    ap.add_argument("--cpu-budget", type=float, default=16.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--form", default="both", choices=["both", "exact", "mfma"],
                    help="Euclidean distance form(s) to time; 'both' reports the MFMA form as value")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the split form's DET scan after the LAM scans instead of beside them")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the untimed extras: decision counters, m = 32 / 64 and fp64-input steps")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (calibrates HBM byte counters)")
    ap.add_argument("--rp-only", action="store_true",
                    help="time only the recurrence plot (the rp extra) and print it as one JSON line "
                         "(for counter passes over the plot kernels)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); a run sharded across GPUs "
              "is not built.", file=sys.stderr)
        sys.exit(2)

    import torch

    import accrqa_amd

    c = workload(args.workload)
    x = np.ascontiguousarray(c["x"], dtype=np.float32)
    tau, emb, thr, lmin, dist = c["tau"][0], c["emb"][0], [float(v) for v in c["thr"]], c["lmin"], c["norms"][0]
    T, nl = len(thr), len(lmin)
    nc = x.size - (emb - 1) * tau

    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    d_rr = torch.zeros(T, dtype=torch.float32, device=dev)
    d_det = torch.zeros(T * nl * 5, dtype=torch.float32, device=dev)
    d_lam = torch.zeros(nl * T * 5, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    L = accrqa_amd.load()
    accrqa_amd.set_overlap(not args.no_overlap)
    if args.copy_probe_mib > 0:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.rp_only:
        print(json.dumps({"rp": rp_bench(accrqa_amd, torch, x, float(c["thr"][5]))}))
        return

    def step():
        accrqa_amd.rqa_device(dx.data_ptr(), x.size, tau, emb, thr, lmin, dist, d_rr.data_ptr(), d_det.data_ptr(),
                              d_lam.data_ptr(), stream=stream.cuda_stream, dtype=np.float32)

    import ctypes as C

    def measure(form):
        accrqa_amd.set_form(form)
        for _ in range(args.warmup):
            step()
        # the timed steps: nothing but the calls between two synchronisations
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = max(ev0.elapsed_time(ev1), wall * 1e3) / args.steps
        outs = (d_rr.cpu().numpy().copy(), d_det.cpu().numpy().copy(), d_lam.cpu().numpy().copy())
        # the same steps with an event pair around every launch, on its stream
        L.accrqa_amd_profile_reset()
        L.accrqa_amd_profile(1)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        wall_ev = time.perf_counter() - w0
        L.accrqa_amd_profile(0)
        kern = {}
        for k in ("k_lam", "k_lam_mx", "k_lam_vscan", "k_lam_hscan", "k_det", "k_det_mx", "k_det_scan", "k_rr",
                  "k_stitch", "k_finalize"):
            t, n = C.c_double(0), C.c_longlong(0)
            L.accrqa_amd_profile_read(k.encode(), C.byref(t), C.byref(n))
            if n.value:
                kern[k] = {"ms_total": t.value, "launches": n.value}
        return ms, kern, outs, wall_ev * 1e3 / args.steps

    forms = {"exact": accrqa_amd.FORM_EXACT, "mfma": accrqa_amd.FORM_MFMA}
    todo = ["exact", "mfma"] if args.form == "both" else [args.form]
    if dist != 1:
        todo = ["exact"]
    res = {f: measure(forms[f]) for f in todo}
    accrqa_amd.set_form(accrqa_amd.FORM_AUTO)
    progress("headline (%s)" % ", ".join(todo))
    head = "mfma" if "mfma" in res else "exact"
    ms_step, kern, outs, _ = res[head]
    mismatches = None
    if "mfma" in res and "exact" in res:
        a, b = res["exact"][2], res["mfma"][2]
        mismatches = int(sum(int((u.view(np.uint32) != v.view(np.uint32)).sum()) for u, v in zip(a, b)))

    value = float(nc) * nc * T / (ms_step / 1e3)
    tiles = any(k in kern for k in ("k_lam_vscan", "k_lam_hscan", "k_det_scan"))   # the split form ran
    form_report = {}
    for f, (ms_f, kern_f, _, ms_ev) in res.items():
        form_report[f] = {"value": float(nc) * nc * T / (ms_f / 1e3), "ms_per_step": ms_f,
                          "ms_per_step_with_kernel_events": ms_ev,
                          "kernels_ms_per_step": {k: v["ms_total"] / args.steps for k, v in kern_f.items()}}
    rooflines = {f: kernel_rooflines(res[f][1], nc, emb, T, dist, args.workload, tiles) for f in res}
    for f in res:
        form_report[f]["rooflines"] = rooflines[f]
    # the dominant kernel of the reported form: the largest on the step's
    # critical path.  With the overlap, k_det_scan runs beside k_lam_vscan and
    # k_lam_hscan; when those two take longer it is off the path.
    crit = dict(kern)
    lam_scans = sum(kern[k]["ms_total"] for k in ("k_lam_vscan", "k_lam_hscan") if k in kern)
    if not args.no_overlap and "k_det_scan" in kern and lam_scans >= kern["k_det_scan"]["ms_total"]:
        del crit["k_det_scan"]
    dom = max(crit, key=lambda k: crit[k]["ms_total"])
    roofline = dict(rooflines[head].get(dom, {}))
    roofline["kernel"] = dom

    rr = d_rr.cpu().numpy()
    out = {
        "metric": "RR+DET+LAM pair-evals/s: MFMA+recheck vs plain FP32 GPU, 0 mismatches vs FP64",
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": c["desc"], "N": int(x.size), "N_c": int(nc), "thresholds": T,
                   "form": {"mfma": "MFMA Gram (FP16 hi/lo split, FP32 accumulation) + exact-order recheck near thresholds",
                            "exact": "exact-order FP32 sums on VALU"}[head],
                   "schedule": ("DET tile scan on a second stream beside the LAM scans" if not args.no_overlap
                                else "tile scans one after the other")},
        "roofline": roofline,
        "kernels_ms_per_step": {k: v["ms_total"] / args.steps for k, v in kern.items()},
        "forms": form_report,
        "mfma_vs_exact_speedup": (form_report["exact"]["ms_per_step"] / form_report["mfma"]["ms_per_step"]
                                  if "mfma" in form_report and "exact" in form_report else None),
        "mismatches_mfma_vs_exact": mismatches,
        "mismatches_note": ("MFMA+recheck form against the exact-order form in the input precision (the "
                            "reference's own decisions), every output bit; the fp32 decisions against true "
                            "fp64 ones are the labelled diagnostic fp32_vs_true_fp64"),
        "rr_min_max": [float(rr.min()), float(rr.max())],
    }
    if not args.no_extras and "mfma" in res:
        out.update(extras(accrqa_amd, L, torch, args, c, x, step))
        if "exact" in res:
            out["gram_precision"] = gram_precision(
                accrqa_amd, L, torch, step,
                lambda: (d_rr.cpu().numpy(), d_det.cpu().numpy(), d_lam.cpu().numpy()), res["exact"][2])
            progress("Gram precision comparison")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(x, tau, emb, thr, lmin, dist, args.cpu_budget)
        progress("cpu baseline")
    print(json.dumps(out), flush=True)


def device_call_seconds(accrqa_amd, torch, x, tau, emb, thr, lmin, dist, steps=3):
    """Seconds per RR+DET+LAM call through the device entry point (series in
    HBM, default form): one warm-up, then the mean of `steps` calls."""
    T, nl = len(thr), len(lmin)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev)
    dx = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(dev)
    o = [torch.zeros(n, dtype=torch.float32, device=dev) for n in (T, T * nl * 5, nl * T * 5)]

    def one():
        accrqa_amd.rqa_device(dx.data_ptr(), x.size, tau, emb, [float(v) for v in thr], lmin, dist, o[0].data_ptr(),
                              o[1].data_ptr(), o[2].data_ptr(), stream=st.cuda_stream, dtype=np.float32)
    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def config1_2(accrqa_amd, torch):
    """The two smaller configurations at their full sizes, through the device
    entry point: config 1 (logistic map, N = 10,000, tau 1, m 3, 5 thresholds;
    the reference's own CPU path timed on the whole of it beside the GPU:
    one warm-up, the median of three runs) and config 2 (Lorenz x, N =
    131,072, tau 8, m 3, 20 thresholds, each norm)."""
    from accrqa_amd import workloads as W
    out = {}
    c = W.config1()
    x = np.ascontiguousarray(c["x"], dtype=np.float32)
    tau, emb, thr, lmin = c["tau"][0], c["emb"][0], [float(v) for v in c["thr"]], c["lmin"]
    nc = x.size - (emb - 1) * tau
    t = device_call_seconds(accrqa_amd, torch, x, tau, emb, thr, lmin, 1, steps=10)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, Reference
    impl, kind = (Reference(), "reference") if Reference.available() else (Oracle(), "port")

    def cpu_run():
        t0 = time.perf_counter()
        impl.RR(x, [tau], [emb], thr, 1)
        impl.DET(x, [tau], [emb], lmin, thr, 1)
        impl.LAM(x, [tau], [emb], lmin, thr, 1)
        return time.perf_counter() - t0
    cpu_run()
    tc = statistics.median([cpu_run() for _ in range(3)])
    out["config1"] = {"n_c": nc, "seconds_per_call": t, "pair_evals_per_s": float(nc) * nc * len(thr) / t,
                      "cpu": {"kind": kind, "cores": Oracle().num_threads(), "seconds_per_call": tc,
                              "pair_evals_per_s": float(nc) * nc * len(thr) / tc,
                              "sample": "the whole configuration, median of 3 runs after a warm-up"}}
    c = W.config2()
    x = np.ascontiguousarray(c["x"], dtype=np.float32)
    tau, emb, thr, lmin = c["tau"][0], c["emb"][0], [float(v) for v in c["thr"]], c["lmin"]
    nc = x.size - (emb - 1) * tau
    out["config2"] = {"n_c": nc}
    for d in c["norms"]:
        t = device_call_seconds(accrqa_amd, torch, x, tau, emb, thr, lmin, d)
        out["config2"]["euclidean" if d == 1 else "maximal"] = {
            "seconds_per_call": t, "pair_evals_per_s": float(nc) * nc * len(thr) / t}
    return out


def config5_step(accrqa_amd, torch, steps=2):
    """The million-sample configuration (Roessler x, N = 1,048,576, tau 1,
    m 10, 32 thresholds, Euclidean, l_min 2): one RR+DET+LAM call through the
    device entry point in the default form (the rank-tile store, 550 GB, does
    not fit whole, so the split form runs slab by slab); one warm-up, then the
    mean of `steps` calls."""
    from accrqa_amd import workloads as W
    t0 = time.perf_counter()
    c = W.config5()
    gen_s = time.perf_counter() - t0
    x = np.ascontiguousarray(c["x"], dtype=np.float32)
    tau, emb, thr, lmin = c["tau"][0], c["emb"][0], [float(v) for v in c["thr"]], c["lmin"]
    T, nl = len(thr), len(lmin)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev)
    dx = torch.from_numpy(x).to(dev)
    o = [torch.zeros(n, dtype=torch.float32, device=dev) for n in (T, T * nl * 5, nl * T * 5)]

    def one():
        accrqa_amd.rqa_device(dx.data_ptr(), x.size, tau, emb, thr, lmin, 1, o[0].data_ptr(), o[1].data_ptr(),
                              o[2].data_ptr(), stream=st.cuda_stream, dtype=np.float32)
    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    nc = x.size - (emb - 1) * tau
    del dx, o
    torch.cuda.empty_cache()
    return {"seconds_per_step": dt, "pair_evals_per_s": float(nc) * nc * T / dt, "n_c": nc, "thresholds": T,
            "generator_seconds": gen_s,
            "path": "accrqa_amd_rqa_device_f32, RR+DET+LAM, default form (split form slab by slab: the "
                    "whole tile store does not fit)"}


def config4_loop(accrqa_amd, stride=64):
    """The batched configuration (4,096 series, tau {1, 2, 4} x m {2, 3, 5} x 8
    thresholds x both norms) through the public host API, one series at a time
    (RR, DET and LAM calls per series and norm), on the systematic sample of
    every stride-th series: the caller's loop that the batched call
    (config4_batched) replaces."""
    from accrqa_amd import workloads as W
    c = W.config4(sample=range(0, 4096, stride))
    pairs = 0
    t0 = time.perf_counter()
    for xs in c["series"]:
        thr = [float(r * xs.std()) for r in c["thr_rel"]]
        for d in c["norms"]:
            accrqa_amd.RR(xs, c["tau"], c["emb"], thr, d)
            accrqa_amd.DET(xs, c["tau"], c["emb"], c["lmin"], thr, d)
            accrqa_amd.LAM(xs, c["tau"], c["emb"], c["lmin"], thr, d)
            for tau in c["tau"]:
                for m in c["emb"]:
                    nc = xs.size - (m - 1) * tau
                    pairs += nc * nc * len(thr)
    dt = time.perf_counter() - t0
    k = len(c["series"])
    return {"sample_series": k, "seconds": dt, "series_per_s": k / dt, "pair_evals_per_s": pairs / dt,
            "full_config_seconds_est": dt * 4096 / k,
            "path": "per-series loop through accrqa_amd.RR/DET/LAM (host arrays), every %dth series" % stride}


def _batch_counts(rep):
    """The batched engine's counters from accrqa_amd.batch_report()."""
    out = {}
    for line in rep.strip().split("\n"):
        if line.startswith("batch:"):
            for part in line[len("batch:"):].split(","):
                v, name = part.strip().split(" ", 1)
                out[name] = int(v)
    return out


def config4_batched(accrqa_amd):
    """The full batched configuration (4,096 series, tau {1, 2, 4} x m {2, 3,
    5} x 8 thresholds at {0.05 .. 0.40} of each series' std, l_min 2) through
    rqa_batch: one call per norm for all series; every block with
    N_c <= 65,535 (here all) through the persistent batched kernels.  A first
    call per norm tunes the kernel's register budget and captures the call's
    launches as a hipGraph (timed and reported as `first_call_seconds`); the
    timed calls (one per norm) replay the graphs.  The series are made (and
    held in host memory) beforehand; every call uploads them.  RR is checked
    against the per-series API for the first series."""
    from accrqa_amd import workloads as W
    t0 = time.perf_counter()
    c = W.config4()
    gen_s = time.perf_counter() - t0
    xs = c["series"]
    thr = np.array([[r * s.std() for r in c["thr_rel"]] for s in xs], dtype=np.float32)
    t0 = time.perf_counter()
    for d in c["norms"]:
        accrqa_amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, d)
    first_s = time.perf_counter() - t0
    before = _batch_counts(accrqa_amd.batch_report())
    t0 = time.perf_counter()
    outs = [accrqa_amd.rqa_batch(xs, c["tau"], c["emb"], c["lmin"], thr, d) for d in c["norms"]]
    dt = time.perf_counter() - t0
    after = _batch_counts(accrqa_amd.batch_report())
    per_call = {k: (after.get(k, 0) - before.get(k, 0)) / len(c["norms"]) for k in after if k != "calls"}
    same = bool(np.array_equal(outs[0]["RR"][0], accrqa_amd.RR(xs[0], c["tau"], c["emb"], thr[0], 1)))
    pairs = 0
    for s in xs:
        for tau in c["tau"]:
            for m in c["emb"]:
                nc = s.size - (m - 1) * tau
                pairs += nc * nc * len(c["thr_rel"]) * len(c["norms"])
    k = len(xs)
    return {"series": k, "seconds": dt, "series_per_s": k / dt, "pair_evals_per_s": pairs / dt,
            "first_call_seconds": first_s, "per_call": per_call,
            "tuning": accrqa_amd.batch_report().strip().split("\n")[:-2],
            "generator_seconds": gen_s, "rr_equals_single_series_api": same,
            "path": "accrqa_amd.rqa_batch, one call per norm over all 4,096 series (persistent batched kernels, "
                    "hipGraph replay)"}


def rp_bench(accrqa_amd, torch, x, thr, reps=3):
    """accrqa_RP through the device entry point, series and plot in HBM, on a
    65,536-row plot (4.3 GB of bytes) of the default workload's series, at
    m = 16 (k_rp_mx_sym, the MFMA form) and m = 3 (k_rp_sym, the exact form),
    both deciding the upper triangle and storing each 64 x 64 square
    mirrored as well, tau = 1, Euclidean; one byte per element written,
    priced against 8 TB/s.  `every_element_form_ms` is the plot with every
    element decided (accrqa_amd_set_rp_sym(0): k_rp_mx, or k_rp4 with four
    adjacent columns per lane), `one_column_form_ms` at m = 3 is k_rp (one
    column per lane, accrqa_amd_set_rp_wide(0)); all with the same plot.
    Beside it, the reference's own accrqa_RP (char output, OpenMP over rows)
    timed on a 6,000-row prefix."""
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    res = {}
    for m in (16, 3):
        nc = 65536
        xs = np.ascontiguousarray(x[:nc + m - 1])
        dx = torch.from_numpy(xs).to(dev)
        dout = torch.empty(nc * nc, dtype=torch.int8, device=dev)
        times, ones = {}, {}
        # (symmetric form, four columns per lane): the default first
        forms = ((1, 1), (0, 1), (0, 0)) if m == 3 else ((1, 1), (0, 1))
        try:
            for form in forms:
                accrqa_amd.set_rp_sym(form[0])
                accrqa_amd.set_rp_wide(form[1])
                accrqa_amd.rp_device(dx.data_ptr(), xs.size, 1, m, thr, 1, dout.data_ptr(), st.cuda_stream)
                torch.cuda.synchronize()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(st)
                for _ in range(reps):
                    accrqa_amd.rp_device(dx.data_ptr(), xs.size, 1, m, thr, 1, dout.data_ptr(), st.cuda_stream)
                b.record(st)
                torch.cuda.synchronize()
                times[form] = a.elapsed_time(b) / reps / 1e3
                ones[form] = int(dout.sum(dtype=torch.int64).item())
        finally:
            accrqa_amd.set_rp_wide(1)
            accrqa_amd.set_rp_sym(1)
        t = times[(1, 1)]
        el = float(nc) * nc
        res[f"m{m}"] = {"n_c": nc, "ms_per_plot": t * 1e3, "elements_per_s": el / t,
                        "write_GB_s": el / t / 1e9, "write_frac_of_8TB_s": el / t / 8e12, "ones": ones[(1, 1)],
                        "every_element_form_ms": times[(0, 1)] * 1e3,
                        "same_plot_ones": len(set(ones.values())) == 1}
        if (0, 0) in times:
            res[f"m{m}"]["one_column_form_ms"] = times[(0, 0)] * 1e3
        del dout, dx
        torch.cuda.empty_cache()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, Reference
    impl, kind = (Reference(), "reference") if Reference.available() else (Oracle(), "port")
    ns = 6000 + 15
    xs = np.ascontiguousarray(x[:ns])
    t0 = time.perf_counter()
    impl.RP(xs, 1, 16, np.float32(thr), 1)
    t = time.perf_counter() - t0
    res["cpu"] = {"kind": kind, "cores": Oracle().num_threads(), "elements_per_s": 6000.0 * 6000 / t,
                  "sample": f"accrqa_RP, N_c = 6,000 prefix, m = 16, {t:.2f} s wall"}
    return res


def rr_target_bench(accrqa_amd, x):
    """RR_target on a 65,536-row prefix of the default series (tau = 1,
    m = 16, target RR 0.05, epsilon 1e-5, so that the search runs all 20
    steps): the speculative bisection (depth 5) against depth 1, which is the
    reference's one RR pass per step, on the same GPU RR."""
    xs = np.ascontiguousarray(x[:65536 + 15])
    res = {"n_c": 65536, "m": 16, "target_RR": 0.05, "epsilon": 1e-5}
    out = {}
    for depth in (5, 1):
        accrqa_amd.RR_target(xs, 1, 16, 0.05, 1, 1e-5, 20, 0, 10, depth=depth)
        t0 = time.perf_counter()
        thr, rr = accrqa_amd.RR_target(xs, 1, 16, 0.05, 1, 1e-5, 20, 0, 10, depth=depth)
        out[depth] = (time.perf_counter() - t0, thr, float(rr))
    res.update({"seconds_speculative": out[5][0], "seconds_sequential": out[1][0],
                "speedup": out[1][0] / out[5][0], "same_result": out[5][1:] == out[1][1:],
                "threshold": out[5][1], "rr": out[5][2]})
    return res


def gram_precision(accrqa_amd, L, torch, step, outputs, exact_outs, steps=2):
    """The measured comparison of Gram precisions on the headline workload:
    the split MFMA form with the rank pass's Gram in FP16 hi/lo (the
    production form), one FP16 half, one BF16 half and one FP8 (E4M3) byte per
    sample, each with its own rigorous error bound.  Per precision: ms per
    step, the rank kernel's time, the elements the table flagged and the
    fraction re-decided in exact order, and the outputs against the exact
    form's bit for bit (every precision must give 0 mismatches)."""
    import ctypes as C
    names = {0: "fp16_hilo", 1: "fp16", 2: "bf16", 3: "fp8_e4m3"}
    out = {}
    accrqa_amd.set_form(accrqa_amd.FORM_MFMA)
    try:
        for g, name in names.items():
            accrqa_amd.set_gram(g)
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            got = outputs()
            mis = int(sum(int((u.view(np.uint32) != v.view(np.uint32)).sum()) for u, v in zip(exact_outs, got)))
            L.accrqa_amd_profile_reset()
            L.accrqa_amd_profile(1)
            step()
            torch.cuda.synchronize()
            L.accrqa_amd_profile(0)
            t, n = C.c_double(0), C.c_longlong(0)
            L.accrqa_amd_profile_read(b"k_lam_mx", C.byref(t), C.byref(n))
            accrqa_amd.mx_stats(True)
            step()
            torch.cuda.synchronize()
            ranked, flagged, exact = accrqa_amd.mx_stats_read()
            accrqa_amd.mx_stats(False)
            out[name] = {"ms_per_step": ms, "k_lam_mx_ms": t.value / max(n.value, 1),
                         "flagged_frac": flagged / max(ranked, 1), "recheck_frac": exact / max(ranked, 1),
                         "mismatches_vs_exact": mis}
    finally:
        accrqa_amd.set_gram(0)
        accrqa_amd.set_form(accrqa_amd.FORM_AUTO)
    return out


def extras(accrqa_amd, L, torch, args, c, x, step):
    """Reported beside the headline, outside its timed region:
    * mfma_decisions: one MFMA-form step with the decision counters on:
      elements ranked by the rank pass, the fraction whose table byte was
      flagged and the fraction re-summed in exact order (the recheck fraction);
    * config3_m: the same series at m = 32 and 64 (thresholds
      linspace(0.1, 1.6) * sqrt(m)), both forms, outputs compared bit for bit;
    * fp64_input: the m = 16 workload with the series and thresholds in fp64,
      both forms (the MFMA form: FP16 hi/lo Gram of the series rounded to FP32, exact
      re-decision in FP64), outputs compared bit for bit;
    * config4_per_series: the batched configuration's throughput through the
      public API, one series at a time (config4_loop);
    * config4_batched: the whole batched configuration through the batched
      API (config4_batched);
    * rp: the recurrence plot (rp_bench);
    * rr_target: the threshold search (rr_target_bench);
    * fp32_vs_true_fp64: the fp32 decisions against true fp64 ones;
    * config5: the million-sample configuration (config5_step);
    * config1, config2: the two smaller configurations (config1_2)."""
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    out = {}
    accrqa_amd.set_form(accrqa_amd.FORM_MFMA)
    accrqa_amd.mx_stats(True)
    step()
    torch.cuda.synchronize()
    ranked, flagged, exact = accrqa_amd.mx_stats_read()
    accrqa_amd.mx_stats(False)
    accrqa_amd.set_form(accrqa_amd.FORM_AUTO)
    out["mfma_decisions"] = {"elements_ranked": ranked, "flagged_frac": flagged / max(ranked, 1),
                             "recheck_frac": exact / max(ranked, 1)}

    def timed(xs, tau, emb, thr, lmin, form, dtype, steps=2):
        T, nl = len(thr), len(lmin)
        dxs = torch.from_numpy(xs).to(dev)
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        o = [torch.zeros(n, dtype=tdt, device=dev) for n in (T, T * nl * 5, nl * T * 5)]
        accrqa_amd.set_form(form)

        def one():
            accrqa_amd.rqa_device(dxs.data_ptr(), xs.size, tau, emb, thr, lmin, 1, o[0].data_ptr(), o[1].data_ptr(),
                                  o[2].data_ptr(), stream=stream.cuda_stream, dtype=dtype)
        one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            one()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        accrqa_amd.set_form(accrqa_amd.FORM_AUTO)
        nc = xs.size - (emb - 1) * tau
        return ms, float(nc) * nc * T / (ms / 1e3), [v.cpu().numpy().copy() for v in o]

    rows = {}
    for m in (32, 64):
        thr = [float(v) for v in np.linspace(0.1, 1.6, 16) * np.sqrt(m)]
        me, ve, oe = timed(x, 1, m, thr, [2], accrqa_amd.FORM_EXACT, np.float32)
        mm, vm, om = timed(x, 1, m, thr, [2], accrqa_amd.FORM_MFMA, np.float32)
        mis = int(sum(int((u.view(np.uint32) != v.view(np.uint32)).sum()) for u, v in zip(oe, om)))
        rows[f"m{m}"] = {"exact_ms_per_step": me, "exact_value": ve, "mfma_ms_per_step": mm, "mfma_value": vm,
                         "mfma_vs_exact_speedup": me / mm, "mismatches_mfma_vs_exact": mis}
    out["config3_m"] = rows
    x64 = np.ascontiguousarray(c["x"], dtype=np.float64)
    args64 = (x64, c["tau"][0], c["emb"][0], [float(v) for v in c["thr"]], c["lmin"])
    me, ve, oe = timed(*args64, accrqa_amd.FORM_EXACT, np.float64)
    mm, vm, om = timed(*args64, accrqa_amd.FORM_MFMA, np.float64)
    mis = int(sum(int((u.view(np.uint64) != v.view(np.uint64)).sum()) for u, v in zip(oe, om)))
    # the reference decides fp32 input in fp32; how many of its decisions a
    # true fp64 evaluation of the same samples would flip (diagnostic)
    t0 = time.perf_counter()
    dx = torch.from_numpy(x).to(dev)
    flips = accrqa_amd.fp64_mismatches(dx.data_ptr(), x.size, c["tau"][0], c["emb"][0], c["thr"], 1,
                                       stream.cuda_stream)
    del dx
    nc = x.size - (c["emb"][0] - 1) * c["tau"][0]
    out["fp32_vs_true_fp64"] = {
        "per_threshold": [int(v) for v in flips], "total": int(flips.sum()),
        "frac_of_decisions": float(flips.sum()) / (float(nc) * nc * len(c["thr"])),
        "seconds": time.perf_counter() - t0,
        "note": ("diagnostic, not a gate: pairs of the full matrix whose decision in the reference's fp32 "
                 "arithmetic differs from a true fp64 evaluation of the same fp32 samples and thresholds "
                 "(k_fp64_diag); the MFMA+recheck form reproduces the fp32 decisions exactly "
                 "(mismatches_mfma_vs_exact), and for fp64 input it matches the exact fp64 form "
                 "(fp64_input.mismatches_mfma_vs_exact)")}
    progress("config3 m=32/64, fp64 input, fp64 diagnostic")
    out["config5"] = config5_step(accrqa_amd, torch)
    progress("config5")
    out.update(config1_2(accrqa_amd, torch))
    progress("config1 and config2")
    out["config4_per_series"] = config4_loop(accrqa_amd)
    progress("config4 per-series sample")
    out["config4_batched"] = config4_batched(accrqa_amd)
    progress("config4 batched")
    out["rp"] = rp_bench(accrqa_amd, torch, x, float(c["thr"][5]))
    out["rr_target"] = rr_target_bench(accrqa_amd, x)
    progress("RP and RR_target")
    out["fp64_input"] = {"exact_ms_per_step": me, "exact_value": ve, "mfma_ms_per_step": mm, "mfma_value": vm,
                         "mfma_vs_exact_speedup": me / mm, "mismatches_mfma_vs_exact": mis,
                         "forms": {"exact": "exact-order FP64 sums on VALU",
                                   "mfma": "FP16 hi/lo MFMA Gram of the series rounded to FP32 + FP64 exact-order recheck"}}
    return out


if __name__ == "__main__":
    main()
