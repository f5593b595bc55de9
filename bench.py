"""Benchmark of the fixed-effect estimation path on one MI355X.

Default workload (the bench line): seconds per estimation call of the 3-FE OLS on the n = 1e7
panel (individual, firm, year; I = 1e6, T = 10, F = 1e5, firm drawn uniformly: the "simple"
graph), inputs already in HBM. A step = one call: structure build (sort, cells, weights) +
residualize [y, x1] + Gram + beta + iid SE, through the C ABI (fh_absorber_create_dev +
fh_ols_fit_dev). value = s/call of the chosen form (FP64 lead + FP32 PCG sweeps refined on FP64
residuals); the plain FP64 form is timed in the same run. The same line carries the difficult
graph (firm = row mod F) and the Poisson PML panel (robust SE), both forms each, with beta / SE
against the oracle. The roofline object is the reduced-operator SpMV S*p, timed live with HIP
events on the absorber's stream. cpu_baseline = the C oracle (OpenMP, every host core) on the
same full-size input, 1 warm-up + median of 3 (the reference itself is Rust and is not built
here).

Other workloads (--workload): "wcr" (K = 10 regressors, CRV1 by firm, WCR11 bootstrap of x1 = 0
with 9999 Rademacher draws; roofline of the int8 MFMA contraction), "ppml" (Poisson PML alone),
"serve" (the mixed request stream with Poisson arrivals: p50/p99 latency and solves/s, against
one request at a time in plain FP64).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0
METRIC = "s/call at n=1e7 3-FE OLS/PPML, mixed vs FP64 GPU at rel Δβ≤1e-8; serve p99"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def line(args, value, ms, dtype, config, roofline, cpu, extra, unit="s/call", higher=False):
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    out["roofline"] = roofline
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def timed(args, fn):
    """fn() over warmup + steps calls; returns (s/call, last result). Python's cyclic garbage
    collector is off inside the timed region, as timeit does: with torch loaded the process
    tracks ~170k objects and a generation-2 pass stops the host for 30-40 ms, which would land on
    whichever call happens to allocate the triggering object (the GPU work is unaffected)."""
    import gc
    import torch
    for _ in range(args.warmup):
        fn()
    gc.collect()
    gc.freeze()
    was = gc.isenabled()
    gc.disable()
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = None
        for _ in range(args.steps):
            res = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, res
    finally:
        if was:
            gc.enable()
        gc.unfreeze()


def measured_traffic(key):
    """HBM bytes per launch measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarized in
    profiles/spmv_ols_n1e7_traffic.md (line "<key>: <bytes>"), or None."""
    path = os.path.join(ROOT, "profiles", "spmv_ols_n1e7_traffic.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def spmv_roofline(F, ab, r, reps, traffic_key=None):
    spmv = {}
    for name, prec in (("fp64", F.FH_PLAIN_FP64), ("fp32", F.FH_MIXED)):
        ms, by = ctypes.c_double(), ctypes.c_double()
        F._lib.check(F.lib().fh_absorber_time_spmv(ab._h, r, prec, reps, ctypes.byref(ms), ctypes.byref(by)))
        spmv[name] = (ms.value, by.value)
        log("spmv r", r, name, "ms", ms.value, "MB", by.value / 1e6, "GB/s", by.value / ms.value / 1e6)
    ms, by = spmv["fp32"]
    achieved = by / (ms * 1e-3) / 1e9
    # the passes' measured ceiling: one L2 gather per cell and pass (the large levels' p in the
    # segment pass, an XCD bucket's segment sums in the level pass), timed by the calibration
    # kernel at the same table sizes and entry width
    info = ab.info
    cells, large = int(info.ncells), int(info.m - info.m_small)
    g1_bucket = max(1, int(info.g1) // 8)
    cal = []
    for levels in (large, g1_bucket):
        t = ctypes.c_double()
        F._lib.check(F.lib().fh_time_gather(max(levels, 1), cells, 4 * r, 20, ctypes.byref(t)))
        cal.append(t.value)
    l2 = {"ceiling_ms": sum(cal), "frac": sum(cal) / ms, "gathers_per_launch": 2 * cells,
          "tables": [large, g1_bucket], "entry_bytes": 4 * r,
          "source": "fh_time_gather: streamed random codes gathering an L2-resident table, one pass per S*p pass"}
    log("spmv l2 gather ceiling ms", sum(cal), "frac", sum(cal) / ms)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "measured_bound": "l2_requests", "l2": l2,
            "traffic": measured_traffic(traffic_key) if traffic_key else None,
            "traffic_source": "profiles/spmv_ols_n1e7_traffic.md (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE, calibrated "
                              "on a 1 GiB device copy)" if traffic_key else None,
            "kernel": f"reduced-operator SpMV S*p (k_small_dense_t + k_seg_t + k_lvl_small + k_bucket_reduce + "
                      f"k_spmv_finish), FP32, r={r}",
            "bytes_per_launch": by, "ms_per_launch": ms,
            "fp64": {"ms": spmv["fp64"][0], "bytes": spmv["fp64"][1],
                     "achieved": spmv["fp64"][1] / (spmv["fp64"][0] * 1e-3) / 1e9}}


def panel_dev(np, torch, synth, n, design, K=1, model="ols", seed=2024):
    d = synth.panel(n, T=10, F=max(1, n // 100), design=design, K=K, model=model, seed=seed)
    dev = torch.device("cuda:0")
    codes_h = [d["indiv"], d["firm"], d["year"]]
    dd = {"codes_h": codes_h, "g": d["g"], "y_h": d["y"], "X_h": d["X"], "firm_h": d["firm"],
          "codes": [torch.as_tensor(c.view(np.int32)).to(dev) for c in codes_h],
          "y": torch.as_tensor(d["y"]).to(dev), "X": torch.as_tensor(d["X"]).to(dev),
          "firm": torch.as_tensor(d["firm"].view(np.int32)).to(dev)}
    torch.cuda.synchronize()
    return dd


def host_cpu():
    """The host's CPU as the CPU baseline reports it: nproc and lscpu's model and topology."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def median_of(fn, reps=3, warmup=1):
    """The reference's timing method (README.md:31-33): warm-up runs, then the median of `reps`."""
    import statistics
    for _ in range(warmup):
        fn()
    ts, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), out


def oracle_ols_call(np, d, K=1, tol=1e-8):
    """One oracle estimation call: structure, residualize, OLS with iid SE (the CPU baseline's
    unit of work, and the parity reference)."""
    from oracle import estimate
    from oracle.oracle import OracleSolver
    s = OracleSolver(d["codes_h"], d["g"])
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(np.concatenate([d["y_h"][:, None], d["X_h"]], 1)[pi])
    s.residualize(rows, rows.shape[1], tol=tol)
    coef, bread, e, kept = estimate.ols_from_block(rows)
    n = rows.shape[0]
    dfk = estimate.df_k("iid", estimate.nparams(K, d["g"]), d["g"], [False] * 3, n_design=K)
    V = estimate.vcov("iid", bread, rows[:, 1:], e, None, n, dfk)
    return coef, np.sqrt(np.diag(V))


def parity(np, res, coef, se):
    return {"rel_dbeta": float(np.max(np.abs(res.coef - coef) / np.abs(coef))),
            "rel_dse": float(np.max(np.abs(res.se - se) / se))}


def time_forms(args, F, d, fit):
    """s/call of the mixed and the plain FP64 form (new absorber per call, inputs in HBM)."""
    out = {}
    for name, prec in (("mixed", F.FH_MIXED), ("fp64", F.FH_PLAIN_FP64)):
        def call(prec=prec):
            ab = F.Absorber(d["codes"], d["g"], precision=prec)
            r = fit(ab)
            ab.close()
            return r
        out[name] = timed(args, call)
    return out


def run_ols(args, np, torch, F, synth):
    n = args.n
    I, Fn = n // 10, max(1, n // 100)
    ols = lambda d: (lambda ab: F.feols(d["y"], d["X"], None, None, vcov="iid", opts=F.SolveOpts(tol=1e-8),
                                        absorber=ab))
    designs = {}
    for design in (("simple", "difficult") if args.design == "simple" else (args.design,)):
        log("generating panel n", n, "F", Fn, design)
        d = panel_dev(np, torch, synth, n, design)
        tf = time_forms(args, F, d, ols(d))
        for name, (t, r) in tf.items():
            log(design, name, "s/call", t, "beta", r.coef, "iters", r.iterations)
        par, cpu = {}, None
        if not args.no_cpu:
            from oracle import oracle as orc
            t_cpu, (coef, se) = median_of(lambda: oracle_ols_call(np, d))
            par = {k: parity(np, r, coef, se) for k, (t, r) in tf.items()}
            cpu = {"value": t_cpu, "unit": "s/call", "cores": orc.threads(), "kind": "port",
                   "sample": f"full workload (n={n}, {design} graph): the C oracle's structure + residualize "
                             f"(tol 1e-8) + OLS iid, OpenMP on {orc.threads()} threads, 1 warm-up + median of 3",
                   "host": host_cpu()}
            log(design, "cpu oracle s/call", t_cpu, "parity", par)
        designs[design] = {"tf": tf, "parity": par, "cpu": cpu, "d": d}
    main_d = designs[args.design]
    ab = F.Absorber(main_d["d"]["codes"], main_d["d"]["g"])
    roof = spmv_roofline(F, ab, 2, args.spmv_reps,
                         "traffic_bytes_fp32_r2" if (args.n == 10_000_000 and args.design == "simple") else None)
    info = ab.info
    ab.close()
    extra = {}
    for design, v in designs.items():
        rm, rf = v["tf"]["mixed"][1], v["tf"]["fp64"][1]
        extra[f"c2_{design}"] = {
            "mixed_s_per_call": v["tf"]["mixed"][0], "fp64_s_per_call": v["tf"]["fp64"][0],
            "pcg_iterations": {k: int(r.iterations) for k, (t, r) in v["tf"].items()},
            "parity_vs_oracle": v["parity"],
            "mixed_vs_fp64": {"rel_dbeta": float(np.max(np.abs(rm.coef - rf.coef) / np.abs(rf.coef))),
                              "rel_dse": float(np.max(np.abs(rm.se - rf.se) / rf.se))}}
    if args.design == "simple" and not args.no_ppml:
        extra["c4_ppml"] = ppml_section(args, np, torch, F, synth)
    config = {"workload": f"ols y~x1 | indiv+firm+year, n={n}, I={I}, T=10, F={Fn}, {args.design} graph, iid",
              "n": n, "fe": [I, Fn, 10], "tol": 1e-8}
    tm, tfp = main_d["tf"]["mixed"], main_d["tf"]["fp64"]
    extra.update({"plain_fp64_s_per_call": tfp[0], "parity_vs_oracle": main_d["parity"],
                  "structure": {"cells": int(info.ncells), "m": int(info.m), "m_small": int(info.m_small)}})
    line(args, tm[0], tm[0] * 1e3, "fp32+fp64", config, roof, main_d["cpu"], extra)


def ppml_section(args, np, torch, F, synth):
    """C4: Poisson PML, robust SE, n = args.n, both forms, parity against the IRLS oracle."""
    n = args.n
    log("generating poisson panel n", n)
    d = panel_dev(np, torch, synth, n, "simple", model="poisson", seed=2026)
    fit = lambda ab: F.fepois(d["y"], d["X"], None, None, vcov="hetero", tol=1e-8, opts=F.SolveOpts(tol=1e-8),
                              absorber=ab)
    tf = time_forms(args, F, d, fit)
    for name, (t, r) in tf.items():
        log("ppml", name, "s/call", t, "beta", r.coef, "irls", r.iterations, "pcg", r.pcg_iterations)
    out = {"mixed_s_per_call": tf["mixed"][0], "fp64_s_per_call": tf["fp64"][0],
           "irls_iterations": {k: int(r.iterations) for k, (t, r) in tf.items()},
           "pcg_iterations": {k: int(r.pcg_iterations) for k, (t, r) in tf.items()}}
    if not args.no_cpu:
        from oracle import poisson
        t0 = time.perf_counter()
        ref = poisson.ppml(d["codes_h"], d["g"], d["y_h"], d["X_h"], tol=1e-8, itol=1e-8, vcov="hetero")
        log("ppml oracle s", time.perf_counter() - t0, "iterations", ref["iterations"])
        out["parity_vs_oracle"] = {k: parity(np, r, ref["coef"], ref["se"]) for k, (t, r) in tf.items()}
        out["oracle_irls_iterations"] = int(ref["iterations"])
    return out


def run_wcr(args, np, torch, F, synth):
    n, K, reps = args.n, 10, args.reps
    log("generating panel n", n, "K", K)
    d = panel_dev(np, torch, synth, n, "simple", K=K, seed=2025)
    R = np.zeros(K)
    R[0] = 1.0
    out = {}

    def call(precision):
        ab = F.Absorber(d["codes"], d["g"], precision=precision)
        res = F.wildboot(d["y"], d["X"], d["firm"], R, r=0.0, reps=reps, seed=42, absorber=ab,
                         opts=F.SolveOpts(tol=1e-8), nested=[0, 1, 0])
        ab.close()
        return res

    for name, prec in (("mixed", F.FH_MIXED), ("fp64", F.FH_PLAIN_FP64)):
        out[name] = timed(args, lambda: call(prec))
        r = out[name][1]
        log(name, "s/call", out[name][0], "stat", r.stat, "p", r.p_val, "ci", r.conf_int, "boot s",
            r.seconds_bootstrap, "mfma ms", r.contraction_ms)
    r = out["mixed"][1]
    tops = r.contraction_ops / (r.contraction_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": tops, "peak": INT8_PEAK_TOPS, "unit": "TOP/s", "frac": tops / INT8_PEAK_TOPS,
            "traffic": None, "kernel": "k_wb_gemm: int8 MFMA draws x clusters x (7 limbs x (2+4k)) contraction",
            "ops_per_launch": r.contraction_ops, "ms_per_launch": r.contraction_ms}
    rm, rf = out["mixed"][1], out["fp64"][1]
    config = {"workload": f"ols y~x1..x10 | indiv+firm+year, n={n}, CRV1 by firm, WCR11 x1=0, B={reps} Rademacher",
              "n": n, "K": K, "G": int(r.n_clusters), "reps": int(r.reps)}
    line(args, out["mixed"][0], out["mixed"][0] * 1e3, "fp32+fp64 (PCG), int8 MFMA (bootstrap)", config, roof, None,
         {"plain_fp64_s_per_call": out["fp64"][0], "bootstrap_s": r.seconds_bootstrap,
          "wcr": {"stat": r.stat, "se": r.se, "p_val": r.p_val, "conf_int": list(r.conf_int),
                  "fe_not_projected": r.fe_not_projected},
          "mixed_vs_fp64": {"rel_dstat": abs(rm.stat - rf.stat) / abs(rf.stat), "rel_dse": abs(rm.se - rf.se) / rf.se,
                            "max_abs_dt_boot": float(np.nanmax(np.abs(rm.t_boot - rf.t_boot)))}})


def run_ppml(args, np, torch, F, synth):
    sec = ppml_section(args, np, torch, F, synth)
    d = panel_dev(np, torch, synth, args.n, "simple", model="poisson", seed=2026)
    ab = F.Absorber(d["codes"], d["g"])
    roof = spmv_roofline(F, ab, 2, args.spmv_reps)
    ab.close()
    config = {"workload": f"poisson count~x1 | indiv+firm+year, n={args.n}, robust SE, tol 1e-8", "n": args.n}
    line(args, sec["mixed_s_per_call"], sec["mixed_s_per_call"] * 1e3, "fp32+fp64", config, roof, None,
         {"plain_fp64_s_per_call": sec["fp64_s_per_call"], "c4_ppml": sec})


def run_serve(args, np, torch, F, synth):
    from fereg_amd import serve
    specs = serve.poisson_stream(args.requests, seed=7)
    log("requests", len(specs), "mean n", float(np.mean([s["n"] for s in specs])), "rate", args.rate)
    res = {}
    for label, slots, prec, pack in (("batched", args.slots, F.FH_MIXED, args.pack),
                                     ("one_at_a_time_fp64", 1, F.FH_PLAIN_FP64, 1)):
        srv = serve.Server(slots=slots, precision=prec, pack=pack)
        w = serve.make_request(specs[0])
        srv.wait(srv.submit(specs[0]["model"], *w))
        replies, wall, late = serve.run_stream(srv, specs, args.rate,
                                               progress=(50, lambda k, lb=label: log(lb, "done", k)))
        srv.close()
        tot = np.array([r.total_s for r in replies])
        res[label] = {"p50_s": float(np.percentile(tot, 50)), "p99_s": float(np.percentile(tot, 99)),
                      "solves_per_s": len(replies) / wall, "errors": int(sum(r.rc != 0 for r in replies)),
                      "max_submit_lateness_s": late,
                      "requests_in_packed_solves": int(sum(r.packed > 1 for r in replies)),
                      "mean_pack": float(np.mean([r.packed for r in replies if r.packed > 1] or [1]))}
        log(label, res[label])
    # Saturated: the first `burst` requests generated beforehand and queued at once, so that
    # solves/s is the engine's own rate (host data generation caps a timed stream near 100/s).
    from concurrent.futures import ThreadPoolExecutor
    bspecs = specs[:args.burst]
    with ThreadPoolExecutor(8) as pool:
        reqs = list(pool.map(serve.make_request, bspecs))
    sat = {}
    for label, slots, prec, pack in (("batched", args.slots, F.FH_MIXED, args.pack),
                                     ("one_at_a_time_fp64", 1, F.FH_PLAIN_FP64, 1)):
        srv = serve.Server(slots=slots, precision=prec, pack=pack)
        srv.wait(srv.submit(bspecs[0]["model"], *reqs[0]))
        replies, wall = serve.run_burst(srv, bspecs, reqs)
        srv.close()
        sat[label] = {"solves_per_s": len(replies) / wall, "wall_s": wall,
                      "errors": int(sum(r.rc != 0 for r in replies)),
                      "mean_service_s": float(np.mean([r.service_s for r in replies])),
                      "requests_in_packed_solves": int(sum(r.packed > 1 for r in replies))}
        log("saturated", label, sat[label])
    del reqs
    res["saturated"] = {"requests": len(bspecs), **sat,
                        "speedup": sat["batched"]["solves_per_s"] / sat["one_at_a_time_fp64"]["solves_per_s"]}
    b = res["batched"]
    config = {"workload": f"serving stream: {len(specs)} requests, Poisson arrivals at {args.rate}/s, n log-uniform "
                          f"[1e4, 2e6], 2 or 3 FEs, K in 1..10, 80% OLS / 20% PPML", "slots": args.slots}
    line(args, b["p99_s"], b["p99_s"] * 1e3, "fp32+fp64", config, None, None,
         {"serve": res, "value_is": "p99 latency (s) of the batched engine"}, unit="s (p99 latency)")


def run_copy(args, np, torch, F, synth):
    """HBM calibration: a plain device copy of a 1 GiB buffer (reads 1 GiB, writes 1 GiB), the
    reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    nbytes = 1 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    s, _ = timed(args, lambda: dst.copy_(src))
    gbs = 2.0 * nbytes / s / 1e9
    log("copy 1 GiB:", s * 1e3, "ms", gbs, "GB/s")
    line(args, gbs, s * 1e3, "u8", {"workload": "device copy, 1 GiB read + 1 GiB written"}, None, None,
         {"bytes_read": nbytes, "bytes_written": nbytes}, unit="GB/s", higher=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--design", default="simple", choices=["simple", "difficult"])
    ap.add_argument("--workload", default="ols", choices=["ols", "wcr", "ppml", "serve", "copy"])
    ap.add_argument("--reps", type=int, default=9999, help="bootstrap draws (wcr)")
    ap.add_argument("--requests", type=int, default=2000, help="requests in the stream (serve)")
    ap.add_argument("--rate", type=float, default=100.0, help="arrivals per second (serve)")
    ap.add_argument("--slots", type=int, default=8, help="requests in flight (serve)")
    ap.add_argument("--pack", type=int, default=32, help="most small OLS requests packed into one solve (serve)")
    ap.add_argument("--burst", type=int, default=400, help="requests queued at once for the saturated rate (serve)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ppml", action="store_true", help="leave the Poisson section out of the ols line")
    ap.add_argument("--spmv-reps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the solver works on one "
              "GPU's share of the design; sharding the FE1 segments over GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fereg_amd as F
    from fereg_amd import synth

    torch.cuda.set_device(0)
    {"ols": run_ols, "wcr": run_wcr, "ppml": run_ppml, "serve": run_serve,
     "copy": run_copy}[args.workload](args, np, torch, F, synth)


if __name__ == "__main__":
    main()
