"""The factor and sample oracle (oracle/sample.py) pinned on the reference's fixtures and on
brute force:
  * fixest's estimation samples (the `obs` rows of tests/data/fit_fixest.json, written by fixest
    0.14.2 from fit_d1.csv / fit_toy.csv): missing values, interactions, singleton removal;
  * reghdfe's e(sample) and num_singletons (tests/data/ols_s*.json, reghdfe 6.13.1 on ols_data.csv):
    the session flavour, frequency-weighted singleton counts;
  * set-based singleton fixpoints and sorted-unique encodings on small random inputs."""
import numpy as np
import pytest

from fixtures import as_list, cases, read_csv
from oracle import sample as S


def _fixest_sample(case, fam):
    d = read_csv(case["data"])
    names = []
    for t in case["fe"]:
        names += [f for f in t["factors"] if f not in names]
    for cl in case["cluster"]:
        names += [f for f in cl if f not in names]
    idx = {nm: i for i, nm in enumerate(names)}
    factors = [(d[nm], "f64", 0, 0) for nm in names]
    return S.build(len(d[case["y"]]), y=d[case["y"]], x=[d[x] for x in as_list(case["x"])],
                   weights=d[case["weights"]] if case.get("weights") else None,
                   offset=d[case["offset"]] if case.get("offset") else None, factors=factors,
                   terms=[tuple(idx[f] for f in t["factors"]) for t in case["fe"]],
                   clusters=[tuple(idx[f] for f in cl) for cl in case["cluster"]], family=fam,
                   separation=fam != "gaussian")


FIXEST_SAMPLES = [(c, "gaussian") for c in cases("fixest_ols.json")] + [(c, "poisson") for c in cases("fixest_pois.json")]


@pytest.mark.parametrize("case,fam", FIXEST_SAMPLES, ids=lambda c: c["name"] if isinstance(c, dict) else c)
def test_fixest_estimation_samples(case, fam):
    r = _fixest_sample(case, fam)
    obs = np.array(case["expect"]["obs"], np.int64) - 1
    np.testing.assert_array_equal(r["kept"], obs)
    assert r["counts"]["nobs"] == case["expect"]["nobs"]


def reghdfe_model(case):
    d = read_csv("ols_data.csv")
    names = []
    for t in case["absorb"] + case["cluster"]:
        names += [f for f in t if f not in names]
    idx = {nm: i for i, nm in enumerate(names)}
    return d, names, idx


@pytest.mark.parametrize("case", cases("reghdfe_ols.json"), ids=lambda c: c["name"])
def test_reghdfe_samples(case):
    d, names, idx = reghdfe_model(case)
    w = case["weight"]
    r = S.build(len(d["y"]), factors=[(d[nm], "f64", 0, 0) for nm in names],
                terms=[tuple(idx[f] for f in t) for t in case["absorb"]],
                clusters=[tuple(idx[f] for f in t) for t in case["cluster"]],
                weights=d[w[0]] if w else None, weight_kind=w[1] if w else None, missing=False,
                singletons=case["singletons"])
    e = case["expect"]
    sample = np.array(e["sample"], float)
    keep = np.zeros(len(sample))
    keep[r["kept"]] = 1.0
    np.testing.assert_array_equal(keep, sample)
    assert r["counts"]["singletons"] == e["scalars"]["num_singletons"]


def _brute_singletons(codes, alive):
    """Set semantics: drop every live row of a level with one live row, over all terms at once,
    until nothing changes (the fixpoint is unique)."""
    alive = alive.copy()
    while True:
        changed = False
        for c in codes:
            cnt = {}
            for i in np.flatnonzero(alive):
                cnt[c[i]] = cnt.get(c[i], 0) + 1
            for i in np.flatnonzero(alive):
                if cnt[c[i]] == 1:
                    alive[i] = False
                    changed = True
        if not changed:
            return alive


@pytest.mark.parametrize("seed", range(6))
def test_singleton_fixpoint_by_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 300
    gs = [int(rng.integers(20, 150)), int(rng.integers(5, 60)), int(rng.integers(2, 40))][: 2 + seed % 2]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    terms = [S.dense_rank(c) for c in codes]
    alive = rng.random(n) > 0.1
    want = _brute_singletons([t[0] for t in terms], alive)
    got = alive.copy()
    s, f = S.drop_singletons(terms, got)
    np.testing.assert_array_equal(got, want)
    assert s == int(alive.sum() - want.sum()) and f == 0


def test_poisson_separation_drops_all_zero_levels():
    # level 0 of the first term has only zero outcomes: its rows go, then the singleton it
    # leaves in the second term
    codes = [np.array([0, 0, 1, 1, 2, 2], np.uint32), np.array([0, 1, 1, 0, 1, 0], np.uint32)]
    y = np.array([0.0, 0.0, 1.0, 2.0, 0.0, 3.0])
    alive = np.ones(6, bool)
    s, f = S.drop_singletons([(codes[0], 3), (codes[1], 2)], alive, y=y, family="poisson")
    assert f == 2 and s == 0 and list(np.flatnonzero(alive)) == [2, 3, 4, 5]
    alive = np.ones(6, bool)
    yl = np.array([0.0, 0.0, 1.0, 1.0, 0.0, 1.0])
    s, f = S.drop_singletons([(codes[0], 3), (codes[1], 2)], alive, y=yl, family="logit")
    # logit also drops the all-one level 1; rows 4 and 5 are then singletons of the second term
    assert f == 4 and s == 2 and not alive.any()


def test_keys_order_values_and_missing_first():
    x = np.array([2.5, -0.0, 0.0, -1e300, np.inf, -np.inf, np.nan, 7.0, -3.25])
    c, g, miss = S.factorize(x)
    # non-finite entries share level 0; the rest follow numeric order; -0.0 == 0.0
    assert miss == 3 and g == 6
    fin = np.isfinite(x)
    assert np.all(c[~fin] == 0)
    order = np.argsort(x[fin], kind="stable")
    assert np.all(np.diff(c[fin][order].astype(int)) >= 0)
    assert c[1] == c[2]
    cs, gs, _ = S.factorize(np.array([np.inf, 1.0, -np.inf, 1.0]), strict=True)
    assert gs == 3 and list(cs) == [2, 1, 0, 1]
    with pytest.raises(ValueError):
        S.factorize(np.array([1.0, np.nan]), strict=True)
    ci, gi, mi = S.factorize(np.array([5, S.INT32_MIN, -7, 5], np.int64), kind="i32")
    assert list(ci) == [2, 0, 1, 2] and gi == 3 and mi == 1
    cc, gc, mc = S.factorize(np.array([3, 9, 4, 1], np.int64), kind="codes", base=1, nlevels=4)
    assert list(cc) == [2, 0, 3, 1] and mc == 1


def test_combine_and_restrict_match_sorted_unique():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 7, 500).astype(np.uint32)
    b = rng.integers(0, 13, 500).astype(np.uint32)
    c, g = S.combine([a, b], [7, 13])
    _, inv = np.unique(np.stack([a, b], 1), axis=0, return_inverse=True)
    np.testing.assert_array_equal(c, inv.reshape(-1))
    kept = np.flatnonzero(rng.random(500) < 0.3)
    r, gr = S.restrict(c, kept)
    _, inv2 = np.unique(c[kept], return_inverse=True)
    np.testing.assert_array_equal(r, inv2.reshape(-1))
    with pytest.raises(ValueError):
        S.combine([a, b, a], [1 << 30, 1 << 30, 1 << 10])
