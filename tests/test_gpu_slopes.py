"""Varying slopes on the GPU (fh_slopes.hip) against the oracle's explicit-design projection
(tests/slopes.rs:97-140 compares the reference's SlopeSolver with the same projection at 1e-9) and
fixest 0.14.2's slope fixtures through the fit entry points."""
import numpy as np
import pytest

from fixtures import cases, slope_inputs
from gpu_util import announce
from oracle import slopes as osl
from test_oracle import fixest_nested
from test_slopes_oracle import panel

pytestmark = pytest.mark.gpu

SPECS = {
    "icpt_firm_slope": [("indiv", True, []), ("year", True, []), ("firm", True, ["x"])],
    "indiv_slope_year": [("indiv", True, ["x"]), ("year", True, [])],
    "firm_slope_only_indiv": [("firm", False, ["x"]), ("indiv", True, [])],
    "no_intercept_term": [("indiv", False, ["x"]), ("firm", False, ["z"])],
    "two_slopes": [("indiv", True, []), ("firm", True, ["x", "z"])],
}


def _design(spec, nindiv=150, nyear=8, nfirm=30, seed=2, weighted=True):
    f, s, w = panel(nindiv, nyear, nfirm, seed, weighted=weighted)
    terms = [(f[t][0], f[t][1], icpt, [s[x] for x in xs]) for t, icpt, xs in spec]
    return terms, w


def _absorber(terms, w, **kw):
    import fereg_amd as F
    return F.Absorber([t[0] for t in terms], [t[1] for t in terms], weights=w,
                      intercepts=[t[2] for t in terms], slopes=[t[3] for t in terms], **kw)


def _columns(terms, n, r, seed):
    """Noise plus every term's intercept and slope effects (tests/slopes.rs:64-88)."""
    rng = np.random.default_rng(seed)
    cols = []
    for _ in range(r):
        v = rng.standard_normal(n) + 5.0
        for codes, g, icpt, xs in terms:
            if icpt:
                v += 3.0 * rng.standard_normal(g)[codes]
            for x in xs:
                v += rng.standard_normal(g)[codes] * x
        cols.append(v)
    return np.stack(cols, axis=1)


def _resid(ab, V, tol=1e-12, warm=None):
    import fereg_amd as F
    pi = ab.row_order()
    rows = np.ascontiguousarray(V[pi])
    info = ab.residualize(rows, V.shape[1], F.SolveOpts(tol=tol, maxiter=10000, want_coefs=True), warm=warm)
    out = np.empty_like(V)
    out[pi] = rows
    return out, info


def _max_rel(a, b):
    return max(np.abs(a[:, j] - b[:, j]).max() / max(np.abs(b[:, j]).max(), 1e-300) for j in range(b.shape[1]))


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("small_limit", [4096, 0])
def test_residuals_match_the_projection(name, small_limit):
    announce("slopes residuals", name, small_limit)
    terms, w = _design(SPECS[name])
    n = len(terms[0][0])
    V = _columns(terms, n, 3, 11)
    ab = _absorber(terms, w, small_limit=small_limit)
    try:
        E, info = _resid(ab, V)
    finally:
        ab.close()
    ref = osl.residualize(V, osl.slope_design(terms, n), w)
    assert _max_rel(E, ref) <= 1e-9, (name, info.iterations)


def test_single_slope_term_is_closed_form():
    terms, w = _design([("indiv", True, ["x", "z"])], nindiv=200, nyear=6, nfirm=10, seed=3)
    n = len(terms[0][0])
    V = _columns(terms, n, 2, 5)
    ab = _absorber(terms, w)
    try:
        E, info = _resid(ab, V)
    finally:
        ab.close()
    assert info.iterations == 1
    np.testing.assert_allclose(E, osl.residualize(V, osl.slope_design(terms, n), w), atol=1e-9 * np.abs(V).max())


@pytest.mark.parametrize("name", ["icpt_firm_slope", "no_intercept_term", "two_slopes"])
def test_operator_matches_the_reduced_matrix(name):
    announce("slopes operator", name)
    terms, w = _design(SPECS[name], nindiv=60, nyear=6, nfirm=12)
    n = len(terms[0][0])
    ab = _absorber(terms, w)
    try:
        bt, bg, boff, _ = ab.blocks()
        m = ab.info.m
        P = np.random.default_rng(3).standard_normal((m, 3))
        Y = ab.apply(P.reshape(-1), 3).reshape(m, 3)
    finally:
        ab.close()
    S = osl.operator(terms, n, w, order=[int(t) for t in bt])
    np.testing.assert_allclose(Y, S @ P, atol=1e-10 * np.abs(S).max() * np.abs(P).max() * m)


def test_weights_and_warm_start():
    """set_weights and a warm start from earlier coefficients against the projection with the new
    weights (rows left out by weight zero, the CRV3 refits, run through the crv3 fixture)."""
    import fereg_amd as F
    terms, w = _design(SPECS["icpt_firm_slope"], nindiv=120, nyear=6, nfirm=40)
    n = len(terms[0][0])
    V = _columns(terms, n, 2, 7)
    ab = _absorber(terms, w)
    try:
        _, info0 = _resid(ab, V)
        rng = np.random.default_rng(61)
        for rnd in range(3):
            wk = 0.3 + 2.0 * rng.uniform(size=n)
            ab.set_weights(np.ascontiguousarray(wk[ab.row_order()]))
            E, info = _resid(ab, V, warm=info0.coefs)
            ref = osl.residualize(V, osl.slope_design(terms, n), wk)
            assert _max_rel(E, ref) <= 1e-9, rnd
    finally:
        ab.close()


def test_fit_rows_and_term_values_reconstruct_the_fit():
    terms, w = _design([("indiv", True, []), ("year", True, []), ("firm", True, ["x", "z"])], nindiv=100)
    n = len(terms[0][0])
    V = _columns(terms, n, 2, 9)
    ab = _absorber(terms, w)
    try:
        E, info = _resid(ab, V)
        pi = ab.row_order()
        for j in range(2):
            fit = np.empty(n)
            fit[pi] = ab.fit_rows(info.coefs, j)
            scale = np.abs(V[:, j]).max()
            np.testing.assert_allclose(V[:, j] - E[:, j], fit, atol=1e-9 * scale)
            vals = ab.term_values(info.coefs, j)
            assert len(vals) == sum(int(t[2]) + len(t[3]) for t in terms)
            recon = np.zeros(n)
            k = 0
            for codes, g, icpt, xs in terms:
                if icpt:
                    recon += vals[k][codes]
                    k += 1
                for x in xs:
                    recon += vals[k][codes] * x
                    k += 1
            np.testing.assert_allclose(recon, fit, atol=1e-9 * scale)
            # normalization: the year intercept has weighted mean 0 (FE1 = indiv takes the shift)
            assert abs(np.sum(w * vals[1][terms[1][0]]) / w.sum()) <= 1e-9 * scale
    finally:
        ab.close()


FIX = cases("fixest_slopes.json")


@pytest.mark.parametrize("case", [c for c in FIX if c["model"] == "ols" and c["vcov"] not in ("hc2", "hc3")],
                         ids=lambda c: c["name"])
def test_fixest_slope_fixtures_ols(case):
    import fereg_amd as F
    announce("slopes fixest", case["name"])
    inp = slope_inputs(case)
    terms = inp["terms"]
    kind = case["vcov"]
    nested = fixest_nested(case, inp) if kind in ("cluster", "crv3") else None
    res = F.feols(inp["y"], inp["X"], [t[0] for t in terms], [t[1] for t in terms], weights=inp["w"], vcov=kind,
                  cluster=inp["clusters"] or None, nested=nested, ssc=case["ssc"], opts=F.SolveOpts(tol=1e-12),
                  intercepts=[t[2] for t in terms], slopes=[t[3] for t in terms])
    ex = case["expect"]
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(ex["coef"]), rtol=1e-8)
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(ex["vcov"], float).reshape(-1), rtol=1e-8)
    assert res.df_k == ex["df_k"] and res.nparams == ex.get("nparams", res.nparams)


@pytest.mark.parametrize("case", [c for c in FIX if c["model"] == "poisson" and c["vcov"] not in ("hc2", "hc3")],
                         ids=lambda c: c["name"])
def test_fixest_slope_fixtures_poisson(case):
    import fereg_amd as F
    announce("slopes fixest poisson", case["name"])
    inp = slope_inputs(case)
    terms = inp["terms"]
    res = F.fepois(inp["y"], inp["X"], [t[0] for t in terms], [t[1] for t in terms], weights=inp["w"],
                   offset=inp["offset"], vcov=case["vcov"], tol=1e-12, opts=F.SolveOpts(tol=1e-12),
                   intercepts=[t[2] for t in terms], slopes=[t[3] for t in terms])
    ex = case["expect"]
    assert res.converged
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(ex["coef"]), rtol=1e-8)
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(ex["vcov"], float).reshape(-1), rtol=1e-7)
    if "deviance" in ex:
        assert abs(res.deviance - ex["deviance"]) <= 1e-8 * ex["deviance"]


def test_what_slope_absorbers_refuse():
    import fereg_amd as F
    terms, w = _design(SPECS["icpt_firm_slope"], nindiv=40, nyear=5, nfirm=8)
    n = len(terms[0][0])
    y = np.random.default_rng(1).standard_normal(n)
    X = np.random.default_rng(2).standard_normal((n, 2))
    with pytest.raises(F.FeregError, match="HC2/HC3"):
        F.feols(y, X, [t[0] for t in terms], [t[1] for t in terms], vcov="hc2",
                intercepts=[t[2] for t in terms], slopes=[t[3] for t in terms])
    with pytest.raises(F.FeregError, match="without intercept or slopes"):
        F.Absorber([terms[0][0]], [terms[0][1]], intercepts=[False])


@pytest.mark.parametrize("weighted", [False, True])
def test_long_levels_run_in_chunks(weighted):
    """Levels longer than the level passes' item size (1024 rows): two year levels of 15 000 rows
    with a slope and a 7-level term; E (50 levels of 600 rows) runs cross the item boundaries."""
    rng = np.random.default_rng(8)
    n = 30_000
    e = np.repeat(np.arange(50), 600).astype(np.uint32)
    year = rng.integers(0, 2, n).astype(np.uint32)
    k7 = rng.integers(0, 7, n).astype(np.uint32)
    x = rng.standard_normal(n)
    w = 0.5 + rng.uniform(size=n) if weighted else None
    terms = [(e, 50, True, []), (year, 2, True, [x]), (k7, 7, True, [])]
    V = _columns(terms, n, 2, 13)
    ab = _absorber(terms, w)
    try:
        E, info = _resid(ab, V)
    finally:
        ab.close()
    ref = osl.residualize(V, osl.slope_design(terms, n), w)
    assert _max_rel(E, ref) <= 1e-9, info.iterations


@pytest.mark.parametrize("weighted", [False, True])
def test_long_segments_of_e_run_in_chunks(weighted):
    """indiv[x] + year: E is year (the only intercept-only term) with two segments of 2000 rows,
    longer than the item size, so its pass runs over items; residuals and coefficients (fit
    reconstruction) against the projection."""
    rng = np.random.default_rng(9)
    n = 4000
    indiv = (np.arange(n) // 10).astype(np.uint32)
    year = (np.arange(n) % 2).astype(np.uint32)
    x = rng.standard_normal(n)
    w = 0.5 + rng.uniform(size=n) if weighted else None
    terms = [(indiv, 400, True, [x]), (year, 2, True, [])]
    V = _columns(terms, n, 3, 17)
    ab = _absorber(terms, w)
    try:
        E, info = _resid(ab, V)
        fit = np.empty(n)
        fit[ab.row_order()] = ab.fit_rows(info.coefs, 1)
    finally:
        ab.close()
    ref = osl.residualize(V, osl.slope_design(terms, n), w)
    assert _max_rel(E, ref) <= 1e-9, info.iterations
    np.testing.assert_allclose(V[:, 1] - E[:, 1], fit, atol=1e-9 * np.abs(V[:, 1]).max())
