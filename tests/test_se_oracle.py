"""The HC2 / HC3 variance of the oracle against fixest 0.14.2's fixtures (tests/data/fit_se_fixest.json
through tests/golden/fixest_se.json; fit/vcov.rs:333-364 for the rule): the leverage is the
exact diagonal of the full model's weighted projection (regressors and every FE dummy), computed
by brute force on the explicit design, so these fixtures pin it."""
import numpy as np
import pytest

from fixtures import case_inputs, cases
from oracle import estimate, poisson
from oracle.oracle import OracleSolver

HC = [c for c in cases("fixest_se.json") if c["vcov"] in ("hc2", "hc3")]


def _expect(case, has_fe):
    e = case["expect"]
    E = np.array(e["vcov"], float).reshape(len(e["coef"]), -1)
    c = np.array(e["coef"], float)
    # without absorbed terms fixest reports the intercept first; the absorber path takes the
    # constant as its only absorbed level, so the regressors' block is compared
    return (c, E) if has_fe else (c[1:], E[1:, 1:])


def ols_hc_oracle(case):
    inp = case_inputs(case)
    codes, gs, y, X, w = inp["codes"], inp["g"], inp["y"], inp["X"], inp["w"]
    n = len(y)
    block = np.concatenate([y[:, None], X], axis=1)
    if codes:
        s = OracleSolver(codes, gs, weights=w)
        pi = s.structure()["pi"]
    else:
        s = OracleSolver([], [], weights=w if w is not None else np.ones(n))
        pi = np.arange(n)
    rows = np.ascontiguousarray(block[pi])
    s.residualize(rows, rows.shape[1], tol=1e-13)
    wr = w[pi] if w is not None else None
    coef, bread, e, kept = estimate.ols_from_block(rows, wr)
    D = estimate.fe_design([c[pi] for c in codes], gs, n, intercept=not codes)
    h = estimate.hat_values(np.concatenate([X[pi][:, kept], D], axis=1), wr, row_weight=bool(codes))
    V = estimate.vcov(case["vcov"], bread, rows[:, 1:][:, kept], e, wr, n, 0.0, h=h)
    return coef, V, h, pi


@pytest.mark.parametrize("case", [c for c in HC if c["model"] == "ols"], ids=lambda c: c["name"])
def test_ols_hc_fixtures(case):
    coef, V, _, _ = ols_hc_oracle(case)
    ec, E = _expect(case, bool(case["fe"]))
    np.testing.assert_allclose(coef, ec, rtol=1e-9)
    np.testing.assert_allclose(V, E, rtol=1e-8, atol=1e-14)


@pytest.mark.parametrize("case", [c for c in HC if c["model"] == "poisson"], ids=lambda c: c["name"])
def test_poisson_hc_fixtures(case):
    inp = case_inputs(case)
    codes, g = inp["codes"], inp["g"]
    n = len(inp["y"])
    if not codes:  # the constant as the one absorbed term
        codes, g = [np.zeros(n, np.uint32)], [1]
    r = poisson.ppml(codes, g, inp["y"], inp["X"], weights=inp["w"], tol=1e-10, itol=1e-12, vcov=case["vcov"])
    ec, E = _expect(case, bool(case["fe"]))
    np.testing.assert_allclose(r["coef"], ec, rtol=1e-9)
    # without absorbed terms fixest's IRLS stops at a slightly different point (coef 1e-11)
    np.testing.assert_allclose(r["vcov"], E, rtol=1e-8 if case["fe"] else 2e-5, atol=1e-14)


def test_leverage_sums_to_the_rank():
    # trace of the projection = rank of the design (FWL split checked on the FE part)
    rng = np.random.default_rng(0)
    n = 300
    codes = [rng.integers(0, 20, n), rng.integers(0, 7, n)]
    X = rng.standard_normal((n, 2))
    w = rng.uniform(0.5, 2.0, n)
    D = estimate.fe_design(codes, [20, 7], n)
    full = np.concatenate([X, D], axis=1)
    h = estimate.hat_values(full, w)
    assert abs(h.sum() - np.linalg.matrix_rank(full)) < 1e-9
    assert np.all(h > 0) and np.all(h < 1)


CRV3 = [c for c in cases("fixest_se.json") if c["vcov"] == "crv3"]


def ols_beta(codes, y, X, w, offset=None):
    """OLS coefficients of y on X with the absorbed terms of codes (the constant without any)."""
    n = len(y)
    gs = [int(c.max()) + 1 for c in codes]
    if codes:
        s = OracleSolver(codes, gs, weights=w)
        pi = s.structure()["pi"]
    else:
        s = OracleSolver([], [], weights=w if w is not None else np.ones(n))
        pi = np.arange(n)
    rows = np.ascontiguousarray(np.concatenate([y[:, None], X], axis=1)[pi])
    s.residualize(rows, rows.shape[1], tol=1e-14)
    return estimate.ols_from_block(rows, None if w is None else w[pi])[0]


def pois_beta(codes, y, X, w, offset=None):
    n = len(y)
    if not codes:
        codes = [np.zeros(n, np.uint32)]
    return poisson.ppml(codes, [int(c.max()) + 1 for c in codes], y, X, weights=w, offset=offset, tol=1e-10,
                        itol=1e-12, vcov="iid")["coef"]


@pytest.mark.parametrize("case", CRV3, ids=lambda c: c["name"])
def test_crv3_fixtures(case):
    inp = case_inputs(case)
    fb = ols_beta if case["model"] == "ols" else pois_beta
    beta, V = estimate.crv3(fb, inp["codes"], inp["y"], inp["X"], inp["w"], inp["clusters"], case["expect"]["df_k"],
                            k_adj=case["ssc"]["K.adj"], g_adj=case["ssc"]["G.adj"], g_df=case["ssc"]["G.df"])
    ec, E = _expect(case, bool(case["fe"]))
    np.testing.assert_allclose(beta, ec, rtol=1e-8)
    np.testing.assert_allclose(V, E, rtol=1e-8, atol=1e-14)
