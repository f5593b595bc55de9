"""The benchmark configurations at their full sizes, HIP path against the oracle.

C1  n = 1e5, indiv + year: plain FP64, iid and CRV1(indiv), the exact path (one step).
C2  n = 1e7, indiv + firm + year, simple and difficult graphs (and a shuffled-row panel):
    the cell structure bit for bit, beta / SE of both forms within 1e-8 of the oracle, and the
    residualized block within the solver tolerance of a tight oracle solve.
C3  n = 1e7, K = 10, CRV1 by firm (G = 1e5), and the WCR11 bootstrap of x1 = 0 against the
    oracle's bootstrap at G = 1e5.
C4  Poisson PML on the C4 panel at n = 1e6, robust SE, both forms against the IRLS oracle.
C5  a request stream drawn from the serving mix: each result the same alone and in the mix.
"""
import functools
import math

import numpy as np
import pytest

from gpu_util import announce
from oracle import estimate
from oracle import poisson as opois
from oracle import wildboot as wb
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu

N_FULL = 10_000_000


@functools.lru_cache(maxsize=3)
def panel(n, design="simple", K=1, model="ols", seed=2024, shuffle=False):
    from fereg_amd import synth
    return synth.panel(n, T=10, F=max(1, n // 100), design=design, K=K, model=model, seed=seed, shuffle=shuffle)


@functools.lru_cache(maxsize=3)
def oracle_ols(n, design, K, seed, shuffle, tol):
    """Oracle structure, residualized block (tol) and the fit's pieces."""
    d = panel(n, design, K, "ols", seed, shuffle)
    codes = [d["indiv"], d["firm"], d["year"]]
    s = OracleSolver(codes, d["g"])
    st = s.structure()
    rows = np.ascontiguousarray(np.concatenate([d["y"][:, None], d["X"]], 1)[st["pi"]])
    s.residualize(rows, rows.shape[1], tol=tol)
    coef, bread, e, kept = estimate.ols_from_block(rows)
    return st, rows, coef, bread, e


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def dev_inputs(d, K=1):
    import torch
    dev = torch.device("cuda:0")
    codes = [torch.as_tensor(d[k].view(np.int32)).to(dev) for k in ("indiv", "firm", "year")]
    return codes, torch.as_tensor(d["y"]).to(dev), torch.as_tensor(d["X"]).to(dev), \
        torch.as_tensor(d["firm"].view(np.int32)).to(dev)


# ---------------------------------------------------------------- C1

@pytest.mark.parametrize("vcov", ["iid", "cluster"])
def test_c1_plain_fp64_matches_oracle(vcov):
    import fereg_amd as F
    d = panel(100_000, seed=11)
    codes, g = [d["indiv"], d["year"]], [d["g"][0], d["g"][2]]
    announce("C1", vcov)
    ab = F.Absorber(codes, g, precision=F.FH_PLAIN_FP64)
    # year (10 levels) is the small block and FE1 = indiv: the dense factor makes the solve exact
    info = ab.residualize(np.ascontiguousarray(np.stack([d["y"], d["X"][:, 0]], 1)[ab.row_order()]), 2)
    assert info.exact and info.iterations == 1
    cl = d["indiv"] if vcov == "cluster" else None
    res = F.feols(d["y"], d["X"], None, None, vcov=vcov, cluster=cl, nested=[True, False] if cl is not None else None,
                  opts=F.SolveOpts(tol=1e-8), absorber=ab)
    ab.close()
    assert res.iterations == 1
    s = OracleSolver(codes, g)
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(np.concatenate([d["y"][:, None], d["X"]], 1)[pi])
    it, exact, _ = s.residualize(rows, 2, tol=1e-8)
    assert exact and it == 1
    coef, bread, e, kept = estimate.ols_from_block(rows)
    n = len(d["y"])
    npar = estimate.nparams(1, g)
    nested = [True, False] if vcov == "cluster" else [False, False]
    dfk = estimate.df_k(vcov, npar, g, nested, n_design=1)
    V = estimate.vcov(vcov, bread, rows[:, 1:], e, None, n, dfk, cluster=None if cl is None else cl[pi])
    assert rel(res.coef, coef) <= 1e-10
    assert rel(res.se, np.sqrt(np.diag(V))) <= 1e-10
    assert abs(res.df_k - dfk) < 1e-9


# ---------------------------------------------------------------- C2

@pytest.mark.parametrize("design,shuffle", [("simple", False), ("difficult", False), ("simple", True)])
def test_c2_structure_bit_exact(design, shuffle):
    import fereg_amd as F
    d = panel(N_FULL, design, 1, "ols", 2024, shuffle)
    codes_d, _, _, _ = dev_inputs(d)
    announce("C2 structure", design, "shuffled" if shuffle else "")
    ab = F.Absorber(codes_d, d["g"])
    sg = ab.structure()
    ab.close()
    so = oracle_ols(N_FULL, design, 1, 2024, shuffle, 1e-10)[0]
    for key in ("pi", "seg_off", "cell_codes"):
        assert np.array_equal(sg[key], so[key]), key
    assert sg["cell_obs_off"] is None and so["cell_obs_off"] is None


@pytest.mark.parametrize("design", ["simple", "difficult"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_c2_fit_matches_oracle(design, precision):
    """beta and the iid SE within 1e-8 of the oracle; the residualized [y, x1] within the solver's
    tolerance of a tight (tol 1e-10) oracle solve."""
    import fereg_amd as F
    import torch
    d = panel(N_FULL, design, 1, "ols", 2024, False)
    codes_d, y_d, X_d, _ = dev_inputs(d)
    st, rows_o, coef, bread, e = oracle_ols(N_FULL, design, 1, 2024, False, 1e-10)
    announce("C2 fit", design, precision)
    ab = F.Absorber(codes_d, d["g"], precision=precision)
    res = F.feols(y_d, X_d, None, None, vcov="iid", opts=F.SolveOpts(tol=1e-8), absorber=ab)
    n = N_FULL
    dfk = estimate.df_k("iid", estimate.nparams(1, d["g"]), d["g"], [False] * 3, n_design=1)
    V = estimate.vcov("iid", bread, rows_o[:, 1:], e, None, n, dfk)
    assert rel(res.coef, coef) <= 1e-8
    assert rel(res.se, np.sqrt(np.diag(V))) <= 1e-8
    # the residualized block itself (tol 1e-8): the FE error is a few tol of the column norms
    pi = torch.as_tensor(st["pi"].astype(np.int64)).to("cuda:0")
    blk = torch.stack([y_d[pi], X_d[pi, 0]], 1).contiguous()
    ab.residualize(blk, 2, F.SolveOpts(tol=1e-8))
    got = blk.cpu().numpy()
    ab.close()
    err = np.sqrt(((got - rows_o) ** 2).sum(0))
    scale = np.sqrt((rows_o ** 2).sum(0))
    assert np.all(err <= 10 * 1e-8 * scale), (err, scale)


# ---------------------------------------------------------------- C3

def _c3():
    return panel(N_FULL, "simple", 10, "ols", 2025, False)


def test_c3_crv1_matches_oracle():
    import fereg_amd as F
    d = _c3()
    codes_d, y_d, X_d, firm_d = dev_inputs(d)
    st, rows_o, coef, bread, e = oracle_ols(N_FULL, "simple", 10, 2025, False, 1e-10)
    announce("C3 CRV1")
    res = F.feols(y_d, X_d, codes_d, d["g"], vcov="cluster", cluster=firm_d, nested=[False, True, False],
                  opts=F.SolveOpts(tol=1e-8))
    g = d["g"]
    dfk = estimate.df_k("cluster", estimate.nparams(10, g), g, [False, True, False], n_design=10)
    V = estimate.vcov("cluster", bread, rows_o[:, 1:], e, None, N_FULL, dfk, cluster=d["firm"][st["pi"]])
    assert rel(res.coef, coef) <= 1e-8
    assert rel(res.se, np.sqrt(np.diag(V))) <= 1e-8
    assert res.n_clusters == g[1]


def test_c3_wcr11_matches_oracle():
    """The bootstrap of x1 = 0 with the firm clusters (G = 1e5): observed statistic, every draw's
    t (999 draws; the oracle holds the draws x clusters sign matrix in host memory), the p-value
    and the test-inverted interval."""
    import fereg_amd as F
    d = _c3()
    codes_d, y_d, X_d, firm_d = dev_inputs(d)
    st, rows_o, coef, bread, e = oracle_ols(N_FULL, "simple", 10, 2025, False, 1e-10)
    R = np.zeros(10)
    R[0] = 1.0
    reps, seed = 999, 42
    announce("C3 WCR11", reps)
    res = F.wildboot(y_d, X_d, firm_d, R, r=0.0, reps=reps, seed=seed, codes=codes_d, g=d["g"],
                     nested=[False, True, False], opts=F.SolveOpts(tol=1e-8))
    g, n, G = d["g"], N_FULL, d["g"][1]
    dfk = estimate.df_k("cluster", estimate.nparams(10, g), g, [False, True, False], n_design=10)
    ssc = (n - 1) / (n - dfk) * G / (G - 1)
    assert abs(res.ssc - ssc) <= 1e-12 * ssc
    assert res.fe_not_projected  # indiv and year cross the firm clusters, G > 512 (mod.rs:992-997)
    ref = wb.wcr11_result(coef, rows_o[:, 1:], e, None, d["firm"][st["pi"]].astype(np.int64), G, bread, R, 0.0, ssc,
                          reps, seed)
    assert abs(res.stat - ref["stat"]) <= 1e-8 * (1 + abs(ref["stat"]))
    assert abs(res.se - ref["se"]) <= 1e-8 * ref["se"]
    np.testing.assert_allclose(res.t_boot, ref["t_boot"], rtol=1e-7, atol=1e-7)
    assert res.p_val == ref["p_val"] and res.reps_feasible == ref["feasible"]
    for a in res.conf_int:
        assert wb.ci_endpoint_ok(a, ref), (res.conf_int, ref["ci"])


# ---------------------------------------------------------------- C4

@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_c4_poisson_matches_oracle(precision):
    import fereg_amd as F
    n = 1_000_000
    d = panel(n, "simple", 1, "poisson", 2026, False)
    codes = [d["indiv"], d["firm"], d["year"]]
    announce("C4 poisson", n, precision)
    ab = F.Absorber(codes, d["g"], precision=precision)
    res = F.fepois(d["y"], d["X"], None, None, vcov="hetero", tol=1e-8, opts=F.SolveOpts(tol=1e-8), absorber=ab)
    ab.close()
    ref = opois.ppml(codes, d["g"], d["y"], d["X"], tol=1e-8, itol=1e-8, vcov="hetero")
    assert res.converged and ref["converged"]
    assert abs(int(res.iterations) - ref["iterations"]) <= 1
    assert rel(res.coef, ref["coef"]) <= 1e-8
    assert rel(res.se, ref["se"]) <= 1e-8


# ---------------------------------------------------------------- C5

def test_c5_stream_results_alone_and_in_the_mix():
    """Requests drawn from the serving mix (n log-uniform in [1e4, 2e6], 2 or 3 FEs, K in 1..10,
    80% OLS / 20% PPML) give the same estimates, bit for bit, run alone or eight at a time."""
    from fereg_amd import serve
    specs = serve.poisson_stream(10, seed=77)
    data = [serve.make_request(s) for s in specs]
    announce("C5 mix", [s["n"] for s in specs])
    srv = serve.Server(slots=8)
    vc = lambda s: "hetero" if s["model"] == "ppml" else "iid"
    tickets = [srv.submit(s["model"], *dd, vcov=vc(s)) for s, dd in zip(specs, data)]
    mixed = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.rc == 0 for r in mixed)
    alone = serve.Server(slots=1)
    for s, dd, rm in zip(specs, data, mixed):
        ra = alone.wait(alone.submit(s["model"], *dd, vcov=vc(s)))
        assert ra.rc == 0
        assert np.array_equal(ra.coef, rm.coef) and np.array_equal(ra.vcov, rm.vcov)
    alone.close()
