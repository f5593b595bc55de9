"""The reghdfe oracle (oracle/reghdfe.py) against reghdfe 6.13.1's ground truth
(tests/data/ols_s*.json, written by tests/data/gen_ols_fixtures.do; the reference holds its own
results to b 1e-7, V 1e-6, the dof table and integer scalars exactly, the other scalars to 1e-6:
tests/ols.rs:843-935)."""
import math

import numpy as np
import pytest

from fixtures import cases, read_csv
from oracle import reghdfe as RH
from oracle import sample as S

INT_SCALARS = ["N", "rank", "df_r", "df_m", "df_a", "df_a_initial", "df_a_redundant", "df_a_nested", "N_clust",
               "N_clustervars", "N_clust1", "N_clust2", "num_singletons", "N_full", "report_constant", "converged"]
FLOAT_SCALARS = ["tss", "tss_within", "rss", "mss", "rmse", "F", "ll", "ll_0", "r2", "r2_within", "r2_a",
                 "r2_a_within", "sumweights"]


def reghdfe_inputs(case):
    """The session sample of a fixture: kept rows, term and cluster codes, weights, y, X."""
    d = read_csv("ols_data.csv")
    names = []
    for t in case["absorb"] + case["cluster"]:
        names += [f for f in t if f not in names]
    idx = {nm: i for i, nm in enumerate(names)}
    w = case["weight"]
    smp = S.build(len(d["y"]), factors=[(d[nm], "f64", 0, 0) for nm in names],
                  terms=[tuple(idx[f] for f in t) for t in case["absorb"]],
                  clusters=[tuple(idx[f] for f in t) for t in case["cluster"]],
                  weights=d[w[0]] if w else None, weight_kind=w[1] if w else None, missing=False,
                  singletons=case["singletons"])
    k = smp["kept"]
    return {"y": d[case["y"]][k], "X": np.stack([d[x][k] for x in case["x"]], 1), "terms": smp["terms"],
            "clusters": [c for c, _ in smp["clusters"]], "w": d[w[0]][k] if w else None,
            "kind": w[1] if w else None, "num_singletons": smp["counts"]["singletons"]}


def vce_of(case):
    return "cluster" if case["cluster"] else ("robust" if case["robust"] else "unadjusted")


def close(a, b, rtol):
    if (a is None or (isinstance(a, float) and math.isnan(a))) and (b is None or (isinstance(b, float) and math.isnan(b))):
        return True
    if a is None or b is None:
        return False
    return abs(a - b) <= rtol * abs(b)


def check_result(case, b, V, dof_tab, scal, status=None):
    e = case["expect"]
    np.testing.assert_allclose(b, np.array(e["b"], float), rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(np.asarray(V).reshape(-1), np.array(e["V"], float), rtol=1e-6, atol=1e-14)
    assert list(map(float, dof_tab)) == list(map(float, e["dof_table"]))
    for name, want in e["scalars"].items():
        if name in INT_SCALARS:
            assert scal[name] == want, name
        elif name in FLOAT_SCALARS:
            got = scal[name]
            assert close(None if got is None or (isinstance(got, float) and math.isnan(got)) else got, want, 1e-6), \
                (name, got, want)
    if case["status"] and status is not None:
        assert list(status) == case["status"]


@pytest.mark.parametrize("case", cases("reghdfe_ols.json"), ids=lambda c: c["name"])
def test_reghdfe_oracle_matches_reghdfe(case):
    inp = reghdfe_inputs(case)
    r = RH.fit(inp["y"], inp["X"], inp["terms"], weights=inp["w"], weight_kind=inp["kind"], vce=vce_of(case),
               clusters=inp["clusters"], noconstant=case["noconstant"], num_singletons=inp["num_singletons"],
               dof_opts=tuple(case["dof"]))
    check_result(case, r["b"], r["V"], r["dof_table"], r["scalars"], r["status"])
    assert r["y_absorbed"] == (case["name"] == "s17_y_absorbed")
