"""OLS and Poisson fits on the GPU against the fixest golden fixtures (fit_fixest.json RTOL 1e-8,
tests/fit_ols.rs:10; GLM b 1e-7 / V 1e-6, tests/glm.rs:3-4)."""
import numpy as np
import pytest

from fixtures import case_inputs, cases
from gpu_util import announce
from test_oracle import fixest_nested

pytestmark = pytest.mark.gpu


def _run(case, model):
    import fereg_amd as F
    inp = case_inputs(case)
    kind = case["vcov"]
    nested = fixest_nested(case, inp) if kind == "cluster" else None
    cl = inp["clusters"][0] if inp["clusters"] else None
    codes, gs = inp["codes"], inp["g"]
    w = inp["w"]
    if not codes and w is None:
        w = np.ones(len(inp["y"]))
    if model == "ols":
        return F.feols(inp["y"], inp["X"], codes, gs, weights=w, vcov=kind, cluster=cl, nested=nested,
                       ssc=case["ssc"], opts=F.SolveOpts(tol=1e-12))
    return F.fepois(inp["y"], inp["X"], codes, gs, vcov=kind, cluster=cl, nested=nested, ssc=case["ssc"],
                    tol=1e-12, opts=F.SolveOpts(tol=1e-12))


@pytest.mark.parametrize("case", [c for c in cases("fixest_ols.json") if len(c["cluster"]) <= 1],
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


@pytest.mark.parametrize("case", [c for c in cases("fixest_pois.json") if not c.get("offset") and not c.get("weights")
                                  and len(c["cluster"]) <= 1], ids=lambda c: c["name"])
def test_fixest_poisson(case):
    announce("fixest poisson", case["name"])
    res = _run(case, "pois")
    e = case["expect"]
    np.testing.assert_allclose(res.coef[res.kept], np.atleast_1d(e["coef"]), rtol=1e-7)
    V = res.vcov[np.ix_(res.kept, res.kept)]
    np.testing.assert_allclose(V.reshape(-1), np.array(e["vcov"], float).reshape(-1), rtol=1e-6)


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
