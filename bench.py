"""Benchmark of the fixed-effect estimation path on one MI355X.

Default workload (the bench line): seconds per estimation call of the 3-FE OLS on the n = 1e7
panel (individual, firm, year; I = 1e6, T = 10, F = 1e5, firm drawn uniformly: the "simple"
graph), inputs already in HBM. A step = one call: structure build (sort, cells, weights) +
residualize [y, x1] + Gram + beta + iid SE, through the C ABI (fh_absorber_create_dev +
fh_ols_fit_dev). value = s/call of the chosen form (FP32 PCG + FP64 refinement); the plain FP64
form is timed in the same run. The roofline object is the reduced-operator SpMV S*p, timed live
with HIP events on the absorber's stream. cpu_baseline = the single-threaded C oracle on the
same full-size input (the reference itself is Rust and is not built here).

Other workloads (--workload): "wcr" (K = 10 regressors, CRV1 by firm, WCR11 bootstrap of x1 = 0
with 9999 Rademacher draws; roofline of the int8 MFMA contraction), "ppml" (Poisson PML, robust
SE), "serve" (the mixed request stream with Poisson arrivals: p50/p99 latency and solves/s,
against one request at a time in plain FP64).
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
    """fn() over warmup + steps calls; returns (s/call, last result)."""
    import torch
    for _ in range(args.warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / args.steps, res


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
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
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


def run_ols(args, np, torch, F, synth):
    n = args.n
    I, Fn = n // 10, max(1, n // 100)
    log("generating panel n", n, "F", Fn, args.design)
    d = panel_dev(np, torch, synth, n, args.design)

    def call(precision):
        ab = F.Absorber(d["codes"], d["g"], precision=precision)
        res = F.feols(d["y"], d["X"], None, None, vcov="iid", opts=F.SolveOpts(tol=1e-8), absorber=ab)
        ab.close()
        return res

    results, timings = {}, {}
    for name, prec in (("mixed", F.FH_MIXED), ("fp64", F.FH_PLAIN_FP64)):
        timings[name], results[name] = timed(args, lambda: call(prec))
        log(name, "s/call", timings[name], "beta", results[name].coef, "iters", results[name].iterations)
    ab = F.Absorber(d["codes"], d["g"])
    roof = spmv_roofline(F, ab, 2, args.spmv_reps,
                         "traffic_bytes_fp32_r2" if (args.n == 10_000_000 and args.design == "simple") else None)
    info = ab.info
    ab.close()
    cpu, dbeta = None, {}
    if not args.no_cpu:
        from oracle import estimate
        from oracle.oracle import OracleSolver
        t0 = time.perf_counter()
        s = OracleSolver(d["codes_h"], d["g"])
        pi = s.structure()["pi"]
        rows = np.ascontiguousarray(np.concatenate([d["y_h"][:, None], d["X_h"]], 1)[pi])
        s.residualize(rows, 2, tol=1e-8)
        coef, bread, e, kept = estimate.ols_from_block(rows)
        V = estimate.vcov("iid", bread, rows[:, 1:], e, None, n, estimate.nparams(1, d["g"]))
        cpu_s = time.perf_counter() - t0
        se_o = np.sqrt(np.diag(V))
        for name in results:
            dbeta[name] = {"rel_dbeta": float(np.max(np.abs(results[name].coef - coef) / np.abs(coef))),
                           "rel_dse": float(np.max(np.abs(results[name].se - se_o) / se_o))}
        cpu = {"value": cpu_s, "unit": "s/call", "cores": 1, "kind": "port",
               "sample": f"full workload (n={n}): oracle structure + residualize (tol 1e-8) + OLS iid, C, one thread"}
        log("cpu oracle s/call", cpu_s, "dbeta", dbeta)
    config = {"workload": f"ols y~x1 | indiv+firm+year, n={n}, I={I}, T=10, F={Fn}, {args.design} graph, iid",
              "n": n, "fe": [I, Fn, 10], "tol": 1e-8}
    rm, rf = results["mixed"], results["fp64"]
    line(args, timings["mixed"], timings["mixed"] * 1e3, "fp32+fp64", config, roof, cpu,
         {"plain_fp64_s_per_call": timings["fp64"],
          "mixed_vs_fp64": {"rel_dbeta": float(np.max(np.abs(rm.coef - rf.coef) / np.abs(rf.coef))),
                            "rel_dse": float(np.max(np.abs(rm.se - rf.se) / rf.se))},
          "pcg_iterations": {k: int(v.iterations) for k, v in results.items()}, "parity_vs_oracle": dbeta,
          "structure": {"cells": int(info.ncells), "m": int(info.m), "m_small": int(info.m_small)}})


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
    n = args.n
    log("generating poisson panel n", n)
    d = panel_dev(np, torch, synth, n, "simple", model="poisson", seed=2026)
    out = {}

    def call(precision):
        ab = F.Absorber(d["codes"], d["g"], precision=precision)
        res = F.fepois(d["y"], d["X"], None, None, vcov="hetero", tol=1e-8, opts=F.SolveOpts(tol=1e-8), absorber=ab)
        ab.close()
        return res

    for name, prec in (("mixed", F.FH_MIXED), ("fp64", F.FH_PLAIN_FP64)):
        out[name] = timed(args, lambda: call(prec))
        log(name, "s/call", out[name][0], "beta", out[name][1].coef, "irls", out[name][1].iterations)
    ab = F.Absorber(d["codes"], d["g"])
    roof = spmv_roofline(F, ab, 2, args.spmv_reps,
                         "traffic_bytes_fp32_r2" if (args.n == 10_000_000 and args.design == "simple") else None)
    ab.close()
    bm, bf = out["mixed"][1], out["fp64"][1]
    config = {"workload": f"poisson count~x1 | indiv+firm+year, n={n}, robust SE, tol 1e-8", "n": n}
    line(args, out["mixed"][0], out["mixed"][0] * 1e3, "fp32+fp64", config, roof, None,
         {"plain_fp64_s_per_call": out["fp64"][0], "irls_iterations": int(bm.iterations),
          "mixed_vs_fp64": {"rel_dbeta": float(np.max(np.abs(bm.coef - bf.coef) / np.abs(bf.coef))),
                            "rel_dse": float(np.max(np.abs(bm.se - bf.se) / bf.se))}})


def run_serve(args, np, torch, F, synth):
    from fereg_amd import serve
    specs = serve.poisson_stream(args.requests, seed=7)
    log("requests", len(specs), "mean n", float(np.mean([s["n"] for s in specs])), "rate", args.rate)
    res = {}
    for label, slots, prec in (("batched", args.slots, F.FH_MIXED), ("one_at_a_time_fp64", 1, F.FH_PLAIN_FP64)):
        srv = serve.Server(slots=slots, precision=prec)
        w = serve.make_request(specs[0])
        srv.wait(srv.submit(specs[0]["model"], *w))
        replies, wall, late = serve.run_stream(srv, specs, args.rate,
                                               progress=(50, lambda k, lb=label: log(lb, "done", k)))
        srv.close()
        tot = np.array([r.total_s for r in replies])
        res[label] = {"p50_s": float(np.percentile(tot, 50)), "p99_s": float(np.percentile(tot, 99)),
                      "solves_per_s": len(replies) / wall, "errors": int(sum(r.rc != 0 for r in replies)),
                      "max_submit_lateness_s": late}
        log(label, res[label])
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
    ap.add_argument("--requests", type=int, default=300, help="requests in the stream (serve)")
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second (serve)")
    ap.add_argument("--slots", type=int, default=8, help="requests in flight (serve)")
    ap.add_argument("--no-cpu", action="store_true")
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
