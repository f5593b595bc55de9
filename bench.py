"""Benchmark: seconds per estimation call of the 3-FE OLS on the n = 1e7 panel (individual,
firm, year; I = 1e6, T = 10, F = 1e5, firm drawn uniformly: the "simple" graph), with the
inputs already in HBM. A step = one call: structure build (sort, cells, weights) + residualize
[y, x1] + Gram + beta + iid SE, through the C ABI (fh_absorber_create_dev + fh_ols_fit_dev).

value = s/call of the chosen form (FP32 PCG + FP64 refinement); the plain FP64 form is timed in
the same run. The roofline object is the reduced-operator SpMV (the PCG hot kernel pair),
timed live with HIP events on the absorber's stream. cpu_baseline = the single-threaded C
oracle on the same full-size input (the reference itself is Rust and is not built here).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "s/call at n=1e7 3-FE OLS/PPML, mixed vs FP64 GPU at rel Δβ≤1e-8; serve p99"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--design", default="simple", choices=["simple", "difficult"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the solver is single-GPU; see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fereg_amd as F
    from fereg_amd import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    n = args.n
    T, I = 10, args.n // 10
    Fn = max(1, I // 10)
    log("generating panel n", n, "F", Fn, args.design)
    d = synth.panel(n, T=T, F=Fn, design=args.design, seed=2024)
    codes_h = [d["indiv"], d["firm"], d["year"]]
    g = d["g"]
    codes = [torch.as_tensor(c.view(np.int32)).to(dev) for c in codes_h]
    y = torch.as_tensor(d["y"]).to(dev)
    X = torch.as_tensor(d["X"]).to(dev)
    torch.cuda.synchronize()

    def call(precision):
        ab = F.Absorber(codes, g, precision=precision)
        res = F.feols(y, X, None, None, vcov="iid", opts=F.SolveOpts(tol=1e-8), absorber=ab)
        ab.close()
        return res

    results = {}
    timings = {}
    for name, prec in (("mixed", F.FH_MIXED), ("fp64", F.FH_PLAIN_FP64)):
        for _ in range(args.warmup):
            call(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            results[name] = call(prec)
        torch.cuda.synchronize()
        timings[name] = (time.perf_counter() - t0) / args.steps
        log(name, "s/call", timings[name], "beta", results[name].coef, "iters", results[name].iterations)

    # SpMV roofline (HIP events on the absorber's stream)
    ab = F.Absorber(codes, g)
    import ctypes
    spmv = {}
    for name, prec in (("fp64", F.FH_PLAIN_FP64), ("fp32", F.FH_MIXED)):
        ms, by = ctypes.c_double(), ctypes.c_double()
        F._lib.check(F.lib().fh_absorber_time_spmv(ab._h, 2, prec, args.spmv_reps, ctypes.byref(ms), ctypes.byref(by)))
        spmv[name] = (ms.value, by.value)
        log("spmv", name, "ms", ms.value, "MB", by.value / 1e6, "GB/s", by.value / ms.value / 1e6)
    info = ab.info
    ab.close()

    # CPU baseline: the oracle on the same input, one thread
    cpu = None
    dbeta = {}
    if not args.no_cpu:
        from oracle import estimate
        from oracle.oracle import OracleSolver
        t0 = time.perf_counter()
        s = OracleSolver(codes_h, g)
        pi = s.structure()["pi"]
        rows = np.ascontiguousarray(np.concatenate([d["y"][:, None], d["X"]], 1)[pi])
        s.residualize(rows, 2, tol=1e-8)
        coef, bread, e, kept = estimate.ols_from_block(rows)
        ssr = float(e @ e)
        V = estimate.vcov("iid", bread, rows[:, 1:], e, None, n, estimate.nparams(1, g))
        cpu_s = time.perf_counter() - t0
        se_o = np.sqrt(np.diag(V))
        for name in results:
            dbeta[name] = {
                "rel_dbeta": float(np.max(np.abs(results[name].coef - coef) / np.abs(coef))),
                "rel_dse": float(np.max(np.abs(results[name].se - se_o) / se_o)),
            }
        cpu = {"value": cpu_s, "unit": "s/call", "cores": 1, "kind": "port",
               "sample": f"full workload (n={n}): oracle structure + residualize (tol 1e-8) + OLS iid, C, one thread"}
        log("cpu oracle s/call", cpu_s, "dbeta", dbeta)

    ms, by = spmv["fp32"]
    achieved = by / (ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": timings["mixed"],
        "unit": "s/call",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": timings["mixed"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"ols y~x1 | indiv+firm+year, n={n}, I={I}, T={T}, F={Fn}, {args.design} graph, iid",
                   "n": n, "fe": [I, Fn, T], "tol": 1e-8},
        "plain_fp64_s_per_call": timings["fp64"],
        "pcg_iterations": {k: int(v.iterations) for k, v in results.items()},
        "parity_vs_oracle": dbeta,
        "structure": {"cells": int(info.ncells), "m": int(info.m), "m_small": int(info.m_small)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "reduced-operator SpMV S*p (k_seg_t + k_small_dense + k_small_dense_final + k_lvl_spmv), FP32, r=2",
                     "bytes_per_launch": by, "ms_per_launch": ms,
                     "fp64": {"ms": spmv["fp64"][0], "bytes": spmv["fp64"][1],
                              "achieved": spmv["fp64"][1] / (spmv["fp64"][0] * 1e-3) / 1e9}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
