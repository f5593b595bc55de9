"""The Poisson IRLS oracle (oracle/poisson.py) against the reference's fixest fixtures
(tests/data/fit_fixest_glm.json, written by fixest 0.14.2 with glm.tol = 1e-12; the reference
holds its own fit to 1e-8 on them, tests/fit_glm.rs:12-34). The fixtures fix the whole IRLS path:
iteration count, deviance, eta and the final IRLS weights, the coefficients and the variance."""
import numpy as np
import pytest

from fixtures import case_inputs, cases
from oracle import poisson
from test_oracle import fixest_nested

RTOL = 1e-8


def _run(case, **kw):
    inp = case_inputs(case)
    nested = fixest_nested(case, inp) if case["vcov"] == "cluster" else None
    cl = inp["clusters"] if inp["clusters"] else None
    return inp, poisson.ppml(inp["codes"], inp["g"], inp["y"], inp["X"], weights=inp["w"], offset=inp["offset"],
                             tol=1e-12, itol=1e-12, vcov=case["vcov"], cluster=cl, nested=nested, **kw)


@pytest.mark.parametrize("case", cases("fixest_pois.json"), ids=lambda c: c["name"])
def test_poisson_oracle_matches_fixest(case):
    inp, r = _run(case)
    e = case["expect"]
    kept = r["kept"]
    assert r["converged"]
    assert r["iterations"] == e["iterations"]
    np.testing.assert_allclose(r["coef"][kept], e["coef"], rtol=RTOL)
    np.testing.assert_allclose(r["vcov"][np.ix_(kept, kept)].ravel(), np.array(e["vcov"], float), rtol=RTOL)
    assert abs(r["deviance"] - e["deviance"]) <= RTOL * abs(e["deviance"])
    assert abs(r["df_k"] - e["df_k"]) < 1e-9
    if e.get("eta"):
        n = len(inp["y"])
        eta = np.empty(n)
        eta[r["pi"]] = r["eta"]
        np.testing.assert_allclose(eta, e["eta"], rtol=0, atol=1e-10)
    if e.get("irls_weights"):
        # fixest reports the weights of the last solve's fill, which the oracle returns as w
        w = np.empty(len(inp["y"]))
        w[r["pi"]] = r["w"]
        np.testing.assert_allclose(w, e["irls_weights"], rtol=1e-8)


def test_poisson_collinear_columns_are_dropped():
    case = next(c for c in cases("fixest_pois.json") if c["name"] == "pois_collin")
    _, r = _run(case)
    assert list(r["kept"]) == [True, True, False, False]
    assert np.all(r["coef"][2:] == 0) and np.all(r["vcov"][2:, :] == 0)


def test_predict_eps_follows_the_history_rule():
    # glm/irls.rs:920-935: NaN until three finite values, then the cubic's extrapolation
    h = [float("nan")] * 3
    assert np.isnan(poisson.predict_eps(h, float("nan")))
    assert np.isnan(poisson.predict_eps(h, float("nan")))
    assert np.isnan(poisson.predict_eps(h, 1e-2))
    assert np.isnan(poisson.predict_eps(h, 1e-4))
    p = poisson.predict_eps(h, 1e-6)
    # ln eps = -2i ln 10 exactly (linear in i): the cubic is that line, extrapolated to 1e-8
    assert abs(p - 1e-8) <= 1e-12


def test_invsym_inverts_and_drops():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((6, 4))
    G = A.T @ A
    np.testing.assert_allclose(poisson.invsym(G), np.linalg.inv(G), rtol=1e-10, atol=1e-12)
    B = np.concatenate([A, A[:, :1] * 2.0], axis=1)  # column 4 = 2 x column 0
    inv = poisson.invsym(B.T @ B)
    assert np.all(inv[4, :] == 0) and np.all(inv[:, 4] == 0)
    np.testing.assert_allclose(inv[:4, :4], np.linalg.inv(G), rtol=1e-8)
