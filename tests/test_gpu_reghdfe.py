"""The GPU reghdfe path (fh_reghdfe_fit through fereg_amd.reghdfe) against reghdfe 6.13.1's
ground truth (tests/data/ols_s*.json: b 1e-7, V 1e-6, the dof table and integer scalars exactly,
the other scalars 1e-6, e(sample) exactly; tests/ols.rs:843-935) and against the oracle on larger
random designs (weights of every kind, two-way clustering, nested terms)."""
import numpy as np
import pytest

from fixtures import cases, read_csv
from gpu_util import announce
from oracle import reghdfe as RH
from oracle import sample as S
from test_reghdfe_oracle import check_result, vce_of

pytestmark = pytest.mark.gpu

DOF_NAMES = ["firstpair", "pairwise", "clusters", "continuous"]


@pytest.mark.parametrize("case", cases("reghdfe_ols.json"), ids=lambda c: c["name"])
def test_reghdfe_fixtures(case):
    from fereg_amd import reghdfe
    d = read_csv("ols_data.csv")
    w = case["weight"]
    announce("reghdfe", case["name"])
    res = reghdfe(d[case["y"]], np.stack([d[x] for x in case["x"]], 1),
                    absorb=[tuple(d[f] for f in t) for t in case["absorb"]], vce=vce_of(case),
                    cluster=[tuple(d[f] for f in t) for t in case["cluster"]], weights=d[w[0]] if w else None,
                    weight_kind=w[1] if w else None, noconstant=case["noconstant"], singletons=case["singletons"],
                    dof=[nm for nm, on in zip(DOF_NAMES, case["dof"]) if on], tol=1e-14)
    e = case["expect"]
    np.testing.assert_array_equal(res.sample.astype(float), np.array(e["sample"], float))
    check_result(case, res.b, res.V, res.dof_table.reshape(-1), res.scalars, res.status)
    assert res.scalars["y_absorbed"] == (1.0 if case["name"] == "s17_y_absorbed" else 0.0)


@pytest.mark.parametrize("kind,vce", [(None, "cluster"), ("aweight", "robust"), ("fweight", "cluster"),
                                      ("pweight", "unadjusted")])
def test_reghdfe_matches_the_oracle_at_scale(kind, vce):
    from fereg_amd import reghdfe
    rng = np.random.default_rng(17)
    n = 200_000
    gi, gf, gt = 30_000, 2_000, 15
    i = rng.integers(0, gi, n)
    f = rng.integers(0, gf, n)
    t = rng.integers(0, gt, n)
    region = f % 40  # firm nested in region
    X = rng.standard_normal((n, 3))
    X[:, 2] = X[:, 0] - 2 * X[:, 1]  # collinear: dropped in stage 2
    y = X[:, 0] - 0.5 * X[:, 1] + rng.standard_normal(gi)[i] + rng.standard_normal(gf)[f] + rng.standard_normal(n)
    w = None if kind is None else (rng.integers(1, 4, n).astype(float) if kind == "fweight" else rng.uniform(0.5, 2, n))
    cl = [(region,), (t,)] if vce == "cluster" else []
    announce("reghdfe at scale", kind, vce)
    res = reghdfe(y, X, absorb=[i.astype(float), f.astype(float), t.astype(float)], vce=vce,
                    cluster=[tuple(c.astype(float) for c in dim) for dim in cl], weights=w, weight_kind=kind,
                    tol=1e-12)
    smp = S.build(n, y=y, x=list(X.T), weights=w, weight_kind=kind,
                  factors=[(i.astype(float), "f64", 0, 0), (f.astype(float), "f64", 0, 0), (t.astype(float), "f64", 0, 0),
                           (region.astype(float), "f64", 0, 0)],
                  terms=[(0,), (1,), (2,)], clusters=[(3,), (2,)] if vce == "cluster" else [])
    k = smp["kept"]
    np.testing.assert_array_equal(np.flatnonzero(res.sample), k)
    want = RH.fit(y[k], X[k], smp["terms"], weights=None if w is None else w[k], weight_kind=kind, vce=vce,
                  clusters=[c for c, _ in smp["clusters"]], tol=1e-12, num_singletons=smp["counts"]["singletons"])
    np.testing.assert_allclose(res.b, want["b"], rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(res.V, want["V"], rtol=1e-6, atol=1e-14)
    assert list(res.dof_table.reshape(-1)) == list(map(float, want["dof_table"]))
    assert list(res.status) == list(want["status"])
    for name in ["N", "df_r", "df_m", "df_a", "df_a_nested", "num_singletons"]:
        assert res.scalars[name] == want["scalars"][name], name
    for name in ["rss", "tss", "tss_within", "r2", "r2_within", "F"]:
        assert abs(res.scalars[name] - want["scalars"][name]) <= 1e-7 * abs(want["scalars"][name]), name
