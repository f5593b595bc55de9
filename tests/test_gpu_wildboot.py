"""WCR11 bootstrap on the GPU (int8 MFMA contraction) against the oracle's cell-score formula
(pinned on the CPU against the brute-force refit) on the same residualized design."""
import math

import numpy as np
import pytest

from fixtures import dense_residualize
from gpu_util import announce
from oracle import wildboot as wb
from fixtures import dummies
from test_wildboot_oracle import fit_fwl, nested_design, non_nested_design, path1_inputs

pytestmark = pytest.mark.gpu


def run_both(y, X, cl, G, codes, gs, w, R, r, reps, seed, pval="two-tailed"):
    import fereg_amd as F
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=seed, pval_type=pval,
                     opts=F.SolveOpts(tol=1e-13), absorber=ab)
    # the oracle on the same rows (solver order), residualized by the dense projection
    yo, Xo, clo = y[pi], X[pi], cl[pi]
    wo = None if w is None else w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, wo)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, np.asarray(R, float), r, res.ssc, reps, seed,
                          pval_type=pval)
    ab.close()
    return res, ref


def check(res, ref, rtol=1e-7):
    assert res.full_enumeration == ref["full"] and res.reps == ref["reps"]
    assert abs(res.point - ref["point"]) <= 1e-9 * (1 + abs(ref["point"]))
    assert abs(res.se - ref["se"]) <= 1e-8 * ref["se"]
    assert abs(res.stat - ref["stat"]) <= 1e-8 * (1 + abs(ref["stat"]))
    np.testing.assert_allclose(res.t_boot, ref["t_boot"], rtol=rtol, atol=rtol)
    assert res.p_val == ref["p_val"] and res.reps_feasible == ref["feasible"]
    # p(x) is a step function: every point meeting the search's stopping rule is an endpoint, and
    # two searches may stop at different such points (within about one draw's step of p)
    for a, b in zip(res.conf_int, ref["ci"]):
        assert math.isfinite(a) == math.isfinite(b), (res.conf_int, ref["ci"])
        assert wb.ci_endpoint_ok(a, ref), (a, res.conf_int, ref["ci"])
        assert a == b or abs(a - b) <= 0.25 * ref["se"], (res.conf_int, ref["ci"])


@pytest.mark.parametrize("weighted", [False, True])
def test_wcr11_nested_fe_matches_oracle(weighted):
    G, per = 40, 3
    y, X, cl, store, w = nested_design(2000, G, per, 3, weighted, seed=3)
    announce("wcr11 nested", weighted)
    res, ref = run_both(y, X, cl, G, [store.astype(np.uint32)], [G * per], w, [1.0, 0.0, 0.0], 0.2, 399, 2024)
    assert not res.fe_not_projected
    check(res, ref)
    # the observed statistic's SE is the CRV1 SE of the fit
    assert abs(res.se - res.fit.se[0]) <= 1e-10 * res.se


def test_wcr11_full_enumeration_and_tails():
    G = 11
    y, X, cl, store, w = nested_design(900, G, 2, 2, False, seed=9)
    for pval in ("two-tailed", "equal-tailed", "greater"):
        announce("wcr11 full enumeration", pval)
        res, ref = run_both(y, X, cl, G, [store.astype(np.uint32)], [G * 2], None, [0.0, 1.0], 0.0, 5000, 7, pval)
        assert res.full_enumeration and res.reps == 1 << G
        check(res, ref)


def test_wcr11_large_non_nested_fe_skips_projection():
    """FE with more levels than 512 crossing 600 clusters: the reference's TooLarge plan."""
    rng = np.random.default_rng(12)
    n, G, gi = 9000, 600, 1500
    cl = rng.integers(0, G, n)
    indiv = rng.integers(0, gi, n)
    X = rng.standard_normal((n, 2))
    y = X @ np.array([0.5, -0.2]) + rng.standard_normal(gi)[indiv] + rng.standard_normal(n)
    announce("wcr11 non-nested")
    res, ref = run_both(y, X, cl, G, [indiv.astype(np.uint32)], [gi], None, [1.0, 0.0], 0.5, 999, 99)
    assert res.fe_not_projected
    check(res, ref)


@pytest.mark.parametrize("with_firm", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_wcr11_non_nested_projection(with_firm, weighted):
    """Path 1 (wildboot/mod.rs:959-1180): a year FE crossing the firm clusters is projected out of
    every draw exactly -- against the oracle's Path 1 and the brute-force refits."""
    import fereg_amd as F
    G, T, reps, r = 30, 6, 199, 0.2
    R = np.array([1.0, 0.0])
    y, X, cl, year, codes, gs, w = non_nested_design(900, G, T, 2, weighted, seed=5 + weighted, with_firm=with_firm)
    announce("wcr11 non-nested path 1", with_firm, weighted)
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=99, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    yo, Xo, clo = y[pi], X[pi], cl[pi]
    wo = None if w is None else w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, wo)
    dnn, dt = path1_inputs(year[pi], T, clo, G, wo, with_firm)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, R, r, res.ssc, reps, 99, dnn=dnn, dtilde=dt)
    check(res, ref)
    D = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    stat, se, tb = wb.brute_t_boot(X, y, w, D, cl, G, R, r, reps, 99, res.ssc)
    np.testing.assert_allclose(res.t_boot, tb, rtol=1e-7, atol=1e-8)


def test_wcr11_non_nested_projection_many_clusters():
    """firm + year, cluster(firm) with 3000 firms: Path 1 at a realistic shape."""
    import fereg_amd as F
    G, T, reps, r = 3000, 10, 999, 0.0
    R = np.array([0.0, 1.0])
    y, X, cl, year, codes, gs, w = non_nested_design(60_000, G, T, 2, False, seed=31)
    announce("wcr11 non-nested path 1, G 3000")
    ab = F.Absorber(codes, gs)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=5, opts=F.SolveOpts(tol=1e-13), absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    yo, Xo, clo = y[pi], X[pi], cl[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, None)
    dnn, dt = path1_inputs(year[pi], T, clo, G, None, True)
    ref = wb.wcr11_result(beta, Xt, u, None, clo, G, bread, R, r, res.ssc, reps, 5, dnn=dnn, dtilde=dt)
    check(res, ref)


@pytest.mark.parametrize("with_firm", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_wcr11_per_cluster_projection(with_firm, weighted):
    """Path 2 (wildboot/mod.rs:1290-1363): a year term with more levels (40) than clusters (12)
    crossing them, projected out of every draw one cluster at a time -- against the oracle's
    Path 2 (one projection per target and cluster) and the brute-force refits."""
    import fereg_amd as F
    G, T, reps, r = 12, 40, 199, 0.5
    R = np.array([0.0, 1.0])
    y, X, cl, year, codes, gs, w = non_nested_design(700, G, T, 2, weighted, seed=17 + weighted, with_firm=with_firm)
    announce("wcr11 per-cluster projection", with_firm, weighted)
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=5, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    yo, Xo, clo = y[pi], X[pi], cl[pi]
    wo = None if w is None else w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, wo)
    D = np.concatenate([dummies(c[pi], g) for c, g in zip(codes, gs)], axis=1)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, R, r, res.ssc, reps, 5, p2_design=D)
    check(res, ref)
    Dn = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    stat, se, tb = wb.brute_t_boot(X, y, w, Dn, cl, G, R, r, reps, 5, res.ssc)
    np.testing.assert_allclose(res.t_boot, tb, rtol=1e-7, atol=1e-8)


def test_wcr11_per_cluster_projection_many_levels():
    """Path 1 with more non-nested levels (30) than fit beside the regressors (60 clusters): Path
    1's algebra with the draws as FP64 products, against the oracle's Path 1 and the refits."""
    import fereg_amd as F
    G, T, reps, r = 60, 30, 499, 0.0
    R = np.array([1.0, -1.0, 0.0])
    y, X, cl, year, codes, gs, w = non_nested_design(3000, G, T, 3, True, seed=41)
    announce("wcr11 per-cluster projection, 30 crossing levels")
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=9, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    yo, Xo, clo, wo = y[pi], X[pi], cl[pi], w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, wo)
    dnn, dt = path1_inputs(year[pi], T, clo, G, wo, True)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, R, r, res.ssc, reps, 9, dnn=dnn, dtilde=dt)
    check(res, ref)
    Dn = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    stat, se, tb = wb.brute_t_boot(X, y, w, Dn, cl, G, R, r, reps, 9, res.ssc)
    np.testing.assert_allclose(res.t_boot, tb, rtol=1e-7, atol=1e-8)


def test_wcr11_per_cluster_projection_at_scale():
    """300 clusters, 400 crossing levels, 100k rows (the reference's Path 2 near its cluster
    limit): against the oracle's Path 1 algebra on the same design (the same projection), with
    the residualization by the C oracle."""
    import fereg_amd as F
    from oracle import estimate
    from oracle.oracle import OracleSolver
    G, T, reps, r = 300, 400, 999, 0.1
    R = np.array([0.0, 1.0])
    y, X, cl, year, codes, gs, w = non_nested_design(100_000, G, T, 2, False, seed=43)
    announce("wcr11 per-cluster projection, G 300, 400 levels")
    ab = F.Absorber(codes, gs)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=21, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    s = OracleSolver(codes, gs)
    assert np.array_equal(s.structure()["pi"], pi)
    rows = np.ascontiguousarray(np.column_stack([y, X])[pi])
    s.residualize(rows, rows.shape[1], tol=1e-13)
    beta, bread, u, kept = estimate.ols_from_block(rows)
    Xt, clo = rows[:, 1:], cl[pi]
    dnn = dummies(year[pi], T)
    cnt = np.bincount(clo, minlength=G).astype(float)
    dt = dnn - (wb.cluster_sums(clo, G, dnn) / cnt[:, None])[clo]  # residuals on the nested firm term
    ref = wb.wcr11_result(beta, Xt, u, None, clo, G, bread, R, r, res.ssc, reps, 21, dnn=dnn, dtilde=dt)
    check(res, ref, rtol=1e-6)


@pytest.mark.parametrize("weighted", [False, True])
def test_wcr11_many_levels_many_clusters(weighted):
    """Path 1 with 2100 clusters and 30 crossing levels (more than fit beside the regressors):
    Path 1's algebra with the draws as FP64 products -- against the oracle's Path 1 and the
    brute-force refits."""
    import fereg_amd as F
    rng = np.random.default_rng(2 + weighted)
    n, G, T, reps, r = 8000, 2100, 30, 199, 0.25
    cl = rng.integers(0, G, n)
    year = rng.integers(0, T, n)
    X = rng.standard_normal((n, 2)) + 0.3 * rng.standard_normal(T)[year][:, None]
    y = X @ np.array([0.5, -0.3]) + rng.standard_normal(T)[year] + rng.standard_normal(n) * (1 + 0.5 * (cl % 3))
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    R = np.array([1.0, 0.0])
    announce("wcr11 many levels, G 2100", weighted)
    codes, gs = [year.astype(np.uint32)], [T]
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=13, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected and res.n_clusters == G
    yo, Xo, clo = y[pi], X[pi], cl[pi]
    wo = None if w is None else w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [year[pi].astype(np.uint32)], gs, wo)
    dnn, dt = path1_inputs(year[pi], T, clo, G, wo, False)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, R, r, res.ssc, reps, 13, dnn=dnn, dtilde=dt)
    check(res, ref)
    stat, se, tb = wb.brute_t_boot(X, y, w, dummies(year, T), cl, G, R, r, reps, 13, res.ssc)
    np.testing.assert_allclose(res.t_boot, tb, rtol=1e-7, atol=1e-8)


def test_wcr11_many_levels_many_clusters_nested_firm():
    """The same with a firm term nested in the 2100 firm clusters: the crossing year dummies are
    residualized on it before their Gram -- against the oracle's Path 1."""
    import fereg_amd as F
    G, T, reps, r = 2100, 30, 499, 0.0
    R = np.array([0.0, 1.0])
    y, X, cl, year, codes, gs, w = non_nested_design(9000, G, T, 2, True, seed=61)
    announce("wcr11 many levels, G 2100, nested firm")
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=r, reps=reps, seed=3, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab)
    ab.close()
    assert not res.fe_not_projected
    yo, Xo, clo, wo = y[pi], X[pi], cl[pi], w[pi]
    Xt, u, beta, bread = fit_fwl(yo, Xo, [c[pi] for c in codes], gs, wo)
    dnn, dt = path1_inputs(year[pi], T, clo, G, wo, True)
    ref = wb.wcr11_result(beta, Xt, u, wo, clo, G, bread, R, r, res.ssc, reps, 3, dnn=dnn, dtilde=dt)
    check(res, ref)


@pytest.mark.parametrize("boot_type", [13, 31, 33])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_mnw_matches_brute_force_refits(boot_type, weighted, cross):
    """MNW 13 / 31 / 33 on the GPU (per-cluster Gram blocks and scores on the device, the jackknife
    algebra per cluster, the WCR11 contraction) against refits by definition: restricted
    leave-one-cluster-out jackknife residuals and per-sample CRV3 variances (tests/wildboot.rs:
    538-708); full enumeration of 2^8 draws, FE nested in the clusters and (cross) a year term
    crossing them (the Path-1 basis)."""
    import fereg_amd as F
    from test_wildboot_oracle import _year_design
    G = 8
    y, X, cl, store, year, w = _year_design(240, G, 3 + boot_type, weighted)
    codes = [store.astype(np.uint32), year.astype(np.uint32)] if cross else [store.astype(np.uint32)]
    gs = [2 * G, 4] if cross else [2 * G]
    announce("mnw", boot_type, weighted, cross)
    ab = F.Absorber(codes, gs, weights=w)
    R = np.array([1.0, 0.0])
    res = F.wildboot(y, X, cl.astype(np.uint32), R, r=0.3, reps=999, seed=7, opts=F.SolveOpts(tol=1e-13),
                     absorber=ab, boot_type=boot_type)
    ab.close()
    assert res.full_enumeration and res.reps == 1 << G and not res.fe_not_projected
    ssc = res.ssc
    if boot_type in (13, 33):
        assert abs(ssc - (G - 1) / G) < 1e-15
    D = np.column_stack([dummies(c, g) for c, g in zip(codes, gs)])
    stat, se, tb = wb.brute_mnw(X, y, w, D, cl, G, R, 0.3, 999, 7, ssc, boot_type)
    assert abs(res.stat - stat) <= 1e-8 * abs(stat) and abs(res.se - se) <= 1e-8 * se
    np.testing.assert_allclose(res.t_boot, tb, rtol=1e-7, atol=1e-8)
    assert 0.0 <= res.p_val <= 1.0 and res.conf_int[0] < res.point < res.conf_int[1]
