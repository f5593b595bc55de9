"""OLS and Poisson fits on the GPU against the fixest golden fixtures (fit_fixest.json RTOL 1e-8,
tests/fit_ols.rs:10; GLM b 1e-7 / V 1e-6, tests/glm.rs:3-4)."""
import numpy as np
import pytest

from fixtures import case_inputs, cases, fixest_ols_cases
from gpu_util import announce
from test_oracle import fixest_nested

pytestmark = pytest.mark.gpu


def _run(case, model):
    import fereg_amd as F
    inp = case_inputs(case)
    kind = case["vcov"]
    nested = fixest_nested(case, inp) if kind == "cluster" else None
    cl = inp["clusters"] if inp["clusters"] else None
    codes, gs = inp["codes"], inp["g"]
    w = inp["w"]
    if model == "ols":
        if not codes and w is None:
            w = np.ones(len(inp["y"]))
        return F.feols(inp["y"], inp["X"], codes, gs, weights=w, vcov=kind, cluster=cl, nested=nested,
                       ssc=case["ssc"], opts=F.SolveOpts(tol=1e-12))
    # the fixtures were written with glm.tol = 1e-12 (tests/fit_glm.rs:1-2)
    return F.fepois(inp["y"], inp["X"], codes, gs, weights=w, offset=inp["offset"], vcov=kind, cluster=cl,
                    nested=nested, ssc=case["ssc"], tol=1e-12, opts=F.SolveOpts(tol=1e-12))


@pytest.mark.parametrize("case", fixest_ols_cases(multiway=True),
                         ids=lambda c: c["name"])
def test_fixest_ols(case):
    announce("fixest ols", case["name"])
    res = _run(case, "ols")
    e = case["expect"]
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(e["coef"]), rtol=1e-8)
    k = int(res.kept.sum())
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(e["vcov"], float).reshape(-1), rtol=1e-8)
    assert abs(res.df_k - e["df_k"]) < 1e-9


@pytest.mark.parametrize("case", [c for c in cases("fixest_pois.json") if len(c["cluster"]) <= 1],
                         ids=lambda c: c["name"])
def test_fixest_poisson(case):
    """fixest's IRLS path: iteration count, deviance, coefficients and variance at the
    reference's own tolerance (tests/fit_glm.rs:12)."""
    announce("fixest poisson", case["name"])
    res = _run(case, "pois")
    e = case["expect"]
    assert res.converged and res.iterations == e["iterations"]
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(e["coef"]), rtol=1e-8)
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(e["vcov"], float).reshape(-1), rtol=1e-8)
    assert abs(res.deviance - e["deviance"]) <= 1e-8 * e["deviance"]


@pytest.mark.parametrize("vcov", ["iid", "hetero", "cluster"])
def test_poisson_matches_oracle_weighted_with_offset(vcov):
    """A three-way design with prior weights and an offset against the numpy IRLS oracle."""
    import fereg_amd as F
    from oracle import poisson
    rng = np.random.default_rng(21)
    n = 40_000
    gs = [4000, 300, 9]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    X = rng.standard_normal((n, 2))
    off = 0.2 * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    fe = sum(0.3 * rng.standard_normal(g)[c] for c, g in zip(codes, gs))
    y = rng.poisson(np.exp(0.1 * X[:, 0] - 0.2 * X[:, 1] + fe + off)).astype(float)
    cl = codes[1] if vcov == "cluster" else None
    nested = [False, True, False] if vcov == "cluster" else None
    announce("poisson vs oracle", vcov)
    res = F.fepois(y, X, codes, gs, weights=w, offset=off, vcov=vcov, cluster=cl, nested=nested, tol=1e-10,
                   opts=F.SolveOpts(tol=1e-12))
    ref = poisson.ppml(codes, gs, y, X, weights=w, offset=off, tol=1e-10, itol=1e-12, vcov=vcov, cluster=cl,
                       nested=nested)
    assert res.iterations == ref["iterations"]
    np.testing.assert_allclose(res.coef, ref["coef"], rtol=1e-8)
    np.testing.assert_allclose(res.se, ref["se"], rtol=1e-7)
    assert abs(res.deviance - ref["deviance"]) <= 1e-9 * ref["deviance"]


def test_prior_weights_survive_a_poisson_fit():
    """A weighted absorber through fepois, then feols: the OLS fit is the weighted one (the
    absorber restores the caller's weights after the IRLS), bit for bit."""
    import fereg_amd as F
    rng = np.random.default_rng(5)
    n = 30_000
    gs = [2000, 50]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    w = rng.uniform(0.5, 2.0, n)
    x = rng.standard_normal((n, 1))
    y = x[:, 0] + rng.standard_normal(n)
    yc = rng.poisson(np.exp(0.2 * x[:, 0])).astype(float)
    announce("prior weights after poisson")
    ab = F.Absorber(codes, gs, weights=w)
    F.fepois(yc, x, None, None, vcov="hetero", tol=1e-10, opts=F.SolveOpts(tol=1e-12), absorber=ab)
    r1 = F.feols(y, x, None, None, vcov="hetero", opts=F.SolveOpts(tol=1e-12), absorber=ab)
    ab.close()
    r0 = F.feols(y, x, codes, gs, weights=w, vcov="hetero", opts=F.SolveOpts(tol=1e-12))
    assert np.array_equal(r0.coef, r1.coef) and np.array_equal(r0.vcov, r1.vcov)


def test_poisson_refit_on_one_absorber_is_bitwise_stable():
    """Two Poisson fits in a row on one absorber: the IRLS weights move into the absorber each
    iteration (the buffer they replace becomes a spare), so the second fit starts from the first
    one's buffers; both must give the same estimates to the last bit."""
    import fereg_amd as F
    rng = np.random.default_rng(11)
    n = 60_000
    gs = [3000, 200, 7]
    codes = [rng.permutation(np.arange(n) % g).astype(np.uint32) for g in gs]
    x = rng.standard_normal(n)
    y = rng.poisson(np.exp(0.3 * x + 0.1)).astype(float)
    ab = F.Absorber(codes, gs)
    announce("poisson refit on one absorber")
    r1 = F.fepois(y, x[:, None], None, None, vcov="hetero", tol=1e-10, opts=F.SolveOpts(tol=1e-12), absorber=ab)
    r2 = F.fepois(y, x[:, None], None, None, vcov="hetero", tol=1e-10, opts=F.SolveOpts(tol=1e-12), absorber=ab)
    assert np.array_equal(r1.coef, r2.coef) and np.array_equal(r1.vcov, r2.vcov)
    # a fresh absorber gives the same fit
    r3 = F.fepois(y, x[:, None], codes, gs, vcov="hetero", tol=1e-10, opts=F.SolveOpts(tol=1e-12))
    np.testing.assert_allclose(r3.coef, r1.coef, rtol=1e-12, atol=0)
    ab.close()


FIT_LAYER_CASES = [(c, "ols") for c in fixest_ols_cases(multiway=True)] + [(c, "poisson") for c in cases("fixest_pois.json")]


@pytest.mark.parametrize("case,model", FIT_LAYER_CASES, ids=lambda c: c["name"] if isinstance(c, dict) else c)
def test_fit_layer_from_raw_columns(case, model):
    """fit::fit on the fixture's raw columns: the GPU sample (missing values, singletons,
    separation) gives fixest's rows, then the fit gives its estimates."""
    import fereg_amd as F
    from fixtures import as_list, read_csv
    d = read_csv(case["data"])
    announce("fit layer", case["name"])
    kw = dict(tol=1e-12) if model == "poisson" else {}
    res = F.fit(d[case["y"]], np.stack([d[x] for x in as_list(case["x"])], 1),
                fe=[tuple(d[f] for f in t["factors"]) for t in case["fe"]], model=model, vcov=case["vcov"],
                cluster=[tuple(d[f] for f in c) for c in case["cluster"]],
                weights=d[case["weights"]] if case.get("weights") else None,
                offset=d[case["offset"]] if case.get("offset") else None, ssc=case["ssc"],
                fe_names=["^".join(t["factors"]) for t in case["fe"]],
                cluster_names=["^".join(c) for c in case["cluster"]], opts=F.SolveOpts(tol=1e-12), **kw)
    e = case["expect"]
    np.testing.assert_array_equal(res.obs, np.array(e["obs"], np.int64) - 1)
    assert res.nobs == e["nobs"]
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(e["coef"]), rtol=1e-8 if model == "ols" else 1e-7)
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(e["vcov"], float).reshape(-1),
                               rtol=1e-8 if model == "ols" else 1e-6)


SE_HC = [c for c in cases("fixest_se.json") if c["vcov"] in ("hc2", "hc3")]


@pytest.mark.parametrize("case", SE_HC, ids=lambda c: c["name"])
def test_fixest_hc2_hc3(case):
    """HC2 / HC3 with the exact leverage of the full model (selected inversion of the joint
    reduced matrix on the GPU path) against fixest 0.14.2 (fit_se_fixest.json); without absorbed
    terms the regressors' block of fixest's matrix (its intercept is the absorber's constant)."""
    import fereg_amd as F
    inp = case_inputs(case)
    codes, gs = inp["codes"], inp["g"]
    announce("hc", case["name"])
    n = len(inp["y"])
    # without absorbed terms the absorber's one level is the constant (it needs the sample size)
    ab = F.Absorber(codes, gs, weights=inp["w"] if inp["w"] is not None or codes else np.ones(n))
    if case["model"] == "ols":
        res = F.feols(inp["y"], inp["X"], None, None, vcov=case["vcov"], opts=F.SolveOpts(tol=1e-13), absorber=ab)
    else:
        res = F.fepois(inp["y"], inp["X"], None, None, vcov=case["vcov"], tol=1e-10, opts=F.SolveOpts(tol=1e-13),
                       absorber=ab)
    ab.close()
    e = case["expect"]
    E = np.array(e["vcov"], float).reshape(len(e["coef"]), -1)
    ec = np.array(e["coef"], float)
    if not codes:
        E, ec = E[1:, 1:], ec[1:]
    np.testing.assert_allclose(res.coef, ec, rtol=1e-8)
    rtol = 1e-8 if (codes or case["model"] == "ols") else 2e-5  # fixest's no-FE IRLS stops elsewhere
    np.testing.assert_allclose(res.vcov[np.ix_(res.kept, res.kept)], E, rtol=rtol, atol=1e-14)


@pytest.mark.parametrize("kind", ["hc2", "hc3"])
@pytest.mark.parametrize("weighted", [False, True])
def test_hc_matches_the_oracle_on_three_terms(kind, weighted):
    """A larger three-term design with repeated cells and a small block: the GPU leverage and
    HC variance against the oracle's brute-force hat matrix."""
    import fereg_amd as F
    from oracle import estimate
    from oracle.oracle import OracleSolver
    rng = np.random.default_rng(21)
    n, gs = 3000, [300, 60, 5]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    codes[0] = rng.permutation(np.repeat(np.arange(300), 10)).astype(np.uint32)  # no singleton level
    X = rng.standard_normal((n, 3))
    y = X @ np.array([1.0, -0.5, 0.2]) + rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    announce("hc oracle", kind, weighted)
    res = F.feols(y, X, codes, gs, weights=w, vcov=kind, opts=F.SolveOpts(tol=1e-13))
    s = OracleSolver(codes, gs, weights=w)
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(np.concatenate([y[:, None], X], axis=1)[pi])
    s.residualize(rows, 4, tol=1e-13)
    wr = w[pi] if w is not None else None
    coef, bread, e, kept = estimate.ols_from_block(rows, wr)
    D = estimate.fe_design([c[pi] for c in codes], gs, n)
    h = estimate.hat_values(np.concatenate([X[pi], D], axis=1), wr)
    V = estimate.vcov(kind, bread, rows[:, 1:], e, wr, n, 0.0, h=h)
    np.testing.assert_allclose(res.coef, coef, rtol=1e-9)
    np.testing.assert_allclose(res.vcov, V, rtol=1e-8, atol=1e-16)


def test_hc_with_a_singleton_level_is_an_error():
    """A level with one observation has leverage 1: fixest refuses HC2/HC3 (fit/vcov.rs:337-351)."""
    import fereg_amd as F
    rng = np.random.default_rng(22)
    n = 397  # levels 0..98 with four rows each, level 99 with one
    codes = [np.concatenate([np.repeat(np.arange(99), 4), [99]]).astype(np.uint32),
             rng.integers(0, 7, n).astype(np.uint32)]
    X = rng.standard_normal((n, 2))
    y = X.sum(1) + rng.standard_normal(n)
    announce("hc singleton")
    with pytest.raises(F.FeregError, match="leverage"):
        F.feols(y, X, codes, [100, 7], vcov="hc2", opts=F.SolveOpts(tol=1e-12))
    r = F.feols(y, X, codes, [100, 7], vcov="hetero", opts=F.SolveOpts(tol=1e-12))
    assert np.all(np.isfinite(r.vcov))


SE_CRV3 = [c for c in cases("fixest_se.json") if c["vcov"] == "crv3"]


@pytest.mark.parametrize("case", SE_CRV3, ids=lambda c: c["name"])
def test_fixest_crv3(case):
    """CRV3: every cluster (of every subset of the dimensions) left out by a zero weight, the
    model re-estimated on the GPU (OLS: one residualization; Poisson: the IRLS), against fixest
    0.14.2 (fit_se_fixest.json)."""
    import fereg_amd as F
    inp = case_inputs(case)
    codes, gs, w = inp["codes"], inp["g"], inp["w"]
    if not codes and w is None:
        w = np.ones(len(inp["y"]))
    nested = fixest_nested(case, inp)
    announce("crv3", case["name"])
    if case["model"] == "ols":
        res = F.feols(inp["y"], inp["X"], codes, gs, weights=w, vcov="crv3", cluster=inp["clusters"], nested=nested,
                      ssc=case["ssc"], opts=F.SolveOpts(tol=1e-13))
    else:
        res = F.fepois(inp["y"], inp["X"], codes, gs, weights=w, vcov="crv3", cluster=inp["clusters"],
                       nested=nested, ssc=case["ssc"], tol=1e-10, opts=F.SolveOpts(tol=1e-13))
    e = case["expect"]
    E = np.array(e["vcov"], float).reshape(len(e["coef"]), -1)
    ec = np.array(e["coef"], float)
    if not codes:
        E, ec = E[1:, 1:], ec[1:]
    np.testing.assert_allclose(res.coef, ec, rtol=1e-8)
    np.testing.assert_allclose(res.vcov[np.ix_(res.kept, res.kept)], E, rtol=1e-7, atol=1e-14)
    assert abs(res.df_k - e["df_k"]) < 1e-9
