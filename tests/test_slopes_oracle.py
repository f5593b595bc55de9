"""The varying-slopes oracle (oracle/slopes.py) pinned on fixest 0.14.2's fixtures
(fit_fixest_fe.json sl_*, fit_fixest_glm.json pois_sl*), a closed form and the projection
identities; no GPU."""
import numpy as np
import pytest

from fixtures import cases, slope_inputs
from oracle import estimate, poisson, slopes
from test_oracle import fixest_nested

CASES = cases("fixest_slopes.json")


def ols_beta(codes, sl, y, X, w, offset=None):
    """beta of a fresh OLS fit with slope terms on a row subset (the CRV3 refits)."""
    terms = [(c, int(c.max()) + 1, t[2], xs) for c, t, xs in zip(codes, ols_beta.terms, sl)]
    rows = slopes.residualize(np.concatenate([y[:, None], X], axis=1), slopes.slope_design(terms, len(y)), w)
    return estimate.ols_from_block(rows, w)[0]


def _ols(case):
    inp = slope_inputs(case)
    terms, y, X, w = inp["terms"], inp["y"], inp["X"], inp["w"]
    n = len(y)
    D = slopes.slope_design(terms, n)
    rows = slopes.residualize(np.concatenate([y[:, None], X], axis=1), D, w)
    coef, bread, e, kept = estimate.ols_from_block(rows, w)
    kind = case["vcov"]
    nested = fixest_nested(case, inp) if kind in ("cluster", "crv3") else [False] * len(terms)
    np_ = slopes.nparams(len(kept), terms)
    dfk = slopes.df_k(kind, np_, terms, nested, n_design=len(kept))
    cl = inp["clusters"] or None
    if kind == "crv3":
        ols_beta.terms = terms
        _, V = estimate.crv3(ols_beta, [t[0] for t in terms], y, X, w, inp["clusters"], dfk,
                             slopes=[t[3] for t in terms])
        return coef, V, dfk, np_, e
    h = None
    if kind in ("hc2", "hc3"):
        # the full model's leverage on the explicit design [X, every FE column] (fit/vcov.rs:333-364)
        h = estimate.hat_values(np.concatenate([X[:, kept], D], axis=1), w)
    V = estimate.vcov(kind, bread, rows[:, 1 + kept], e, w, n, dfk, cluster=cl, h=h)
    return coef, V, dfk, np_, e


@pytest.mark.parametrize("case", [c for c in CASES if c["model"] == "ols"], ids=lambda c: c["name"])
def test_ols_slope_fixtures(case):
    coef, V, dfk, np_, e = _ols(case)
    ex = case["expect"]
    np.testing.assert_allclose(coef, np.atleast_1d(ex["coef"]), rtol=1e-8)
    np.testing.assert_allclose(V.reshape(-1), np.array(ex["vcov"], float).reshape(-1), rtol=1e-8)
    assert dfk == ex["df_k"] and np_ == ex.get("nparams", np_)
    if ex.get("resid"):
        np.testing.assert_allclose(e, np.array(ex["resid"]), atol=1e-9 * np.abs(ex["resid"]).max())


@pytest.mark.parametrize("case", [c for c in CASES if c["model"] == "poisson"], ids=lambda c: c["name"])
def test_poisson_slope_fixtures(case):
    inp = slope_inputs(case)
    res = poisson.ppml(None, None, inp["y"], inp["X"], weights=inp["w"], offset=inp["offset"], tol=1e-12,
                       vcov=case["vcov"], slope_terms=inp["terms"])
    ex = case["expect"]
    assert res["converged"]
    if "iterations" in ex and case["name"] != "pois_sl_only":
        # the IRLS count follows the inner solves' accuracy schedule; with the slope-only term the
        # reference's inexact PCG solves take a different path (32 there, 40 with exact
        # projections) to the same estimates
        assert res["iterations"] == ex["iterations"]
    np.testing.assert_allclose(res["coef"], np.atleast_1d(ex["coef"]), rtol=1e-8)
    np.testing.assert_allclose(res["vcov"].reshape(-1), np.array(ex["vcov"], float).reshape(-1), rtol=1e-7)
    if "deviance" in ex:
        assert abs(res["deviance"] - ex["deviance"]) <= 1e-8 * ex["deviance"]
    assert res["df_k"] == ex["df_k"]


def panel(nindiv, nyear, nfirm, seed, weighted=False):
    """indiv x year panel with a random firm, slope x constant within every 7th indiv (no within
    variation: the slope of that level is not identified) and z (tests/slopes.rs:28-41)."""
    rng = np.random.default_rng(seed)
    n = nindiv * nyear
    indiv = np.arange(n) // nyear
    year = np.arange(n) % nyear
    firm = rng.integers(0, nfirm, n)
    x = np.where(indiv % 7 == 0, 1.5, 2.0 + rng.standard_normal(n))
    z = rng.standard_normal(n)
    w = 0.5 + rng.uniform(size=n) if weighted else None
    f = {"indiv": (indiv.astype(np.uint32), nindiv), "year": (year.astype(np.uint32), nyear),
         "firm": (firm.astype(np.uint32), nfirm)}
    return f, {"x": x, "z": z}, w


def test_single_slope_term_is_per_level_regression():
    """One term indiv##c.x: the projection is a weighted regression on (1, x) within each level."""
    f, s, w = panel(60, 6, 5, 3, weighted=True)
    codes, g = f["indiv"]
    v = np.random.default_rng(4).standard_normal(len(codes))
    e = slopes.residualize(v, slopes.slope_design([(codes, g, True, [s["x"]])], len(codes)), w)
    for lv in range(g):
        m = codes == lv
        A = np.stack([np.ones(m.sum()), s["x"][m]], axis=1)
        sw = np.sqrt(w[m])
        b, *_ = np.linalg.lstsq(A * sw[:, None], v[m] * sw, rcond=None)
        np.testing.assert_allclose(e[m], v[m] - A @ b, atol=1e-12)


@pytest.mark.parametrize("spec", [
    [("indiv", True, []), ("year", True, []), ("firm", True, ["x"])],
    [("indiv", True, ["x"]), ("year", True, [])],
    [("firm", False, ["x"]), ("indiv", True, [])],
    [("indiv", False, ["x"]), ("firm", False, ["z"])],
])
def test_projection_identities(spec):
    f, s, w = panel(80, 8, 20, 5, weighted=True)
    terms = [(f[t][0], f[t][1], icpt, [s[x] for x in xs]) for t, icpt, xs in spec]
    n = len(f["indiv"][0])
    D = slopes.slope_design(terms, n)
    V = np.random.default_rng(6).standard_normal((n, 3))
    E = slopes.residualize(V, D, w)
    # D'W e = 0 and the projection is idempotent
    assert np.abs(D.T @ (w[:, None] * E)).max() <= 1e-9 * np.abs(V).max() * n
    np.testing.assert_allclose(slopes.residualize(E, D, w), E, atol=1e-10)
    # S (the reduced matrix after eliminating E) is symmetric PSD, and its range gives the
    # same residuals: rhs = D_R'W (v - H_E t), x = S^+ rhs, e = v - D_R x - H_E t'
    e_idx = slopes.elimination_index(terms)
    S = slopes.operator(terms, n, w)
    np.testing.assert_allclose(S, S.T, atol=1e-10 * np.abs(S).max())
    assert np.linalg.eigvalsh(S).min() >= -1e-9 * np.abs(S).max()
    order = [q for q in range(len(terms)) if q != e_idx]
    DR = slopes.slope_design([terms[q] for q in order], n)
    HE = slopes.slope_design([terms[e_idx]], n)
    PE = lambda A: slopes.residualize(A, HE, w)
    rhs = DR.T @ (w[:, None] * PE(V))
    x = np.linalg.pinv(S, rcond=1e-11, hermitian=True) @ rhs
    np.testing.assert_allclose(PE(V - DR @ x), E, atol=1e-8)


def test_elimination_choice():
    f, s, _ = panel(30, 4, 9, 1)
    t = lambda k, icpt, xs: (f[k][0], f[k][1], icpt, [s[x] for x in xs])
    assert slopes.elimination_index([t("year", True, []), t("indiv", True, ["x"]), t("firm", True, [])]) == 2
    assert slopes.elimination_index([t("firm", False, ["x"]), t("indiv", False, ["z"])]) == 1
    assert slopes.nparams(1, [t("indiv", True, ["x"]), t("year", True, [])]) == 1 + 30 * 2 + 4 - 1
