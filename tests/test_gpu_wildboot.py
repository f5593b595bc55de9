"""WCR11 bootstrap on the GPU (int8 MFMA contraction) against the oracle's cell-score formula
(pinned on the CPU against the brute-force refit) on the same residualized design."""
import math

import numpy as np
import pytest

from fixtures import dense_residualize
from gpu_util import announce
from oracle import wildboot as wb
from test_wildboot_oracle import fit_fwl, nested_design

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


def test_wcr11_small_non_nested_design_is_refused():
    import fereg_amd as F
    rng = np.random.default_rng(2)
    n, G = 500, 30
    cl = rng.integers(0, G, n).astype(np.uint32)
    year = rng.integers(0, 5, n).astype(np.uint32)
    X = rng.standard_normal((n, 1))
    y = X[:, 0] + rng.standard_normal(n)
    with pytest.raises(F.FeregError) as e:
        F.wildboot(y, X, cl, [1.0], reps=99, codes=[year], g=[5])
    assert e.value.rc == 198
