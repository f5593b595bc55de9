"""The GPU factor layer and estimation sample (fh_factorize / fh_combine / fh_restrict /
fh_drop_singletons / fh_sample_*) against the oracle (oracle/sample.py), bit-exact: codes, level
counts, kept rows and drop counts. Both ranking paths are covered (presence table for key ranges
below max(2n, 2^20), radix sort above), as are missing entries, the session flavour and the
reference's fixtures (fixest's estimation samples, reghdfe's e(sample))."""
import numpy as np
import pytest

from fixtures import as_list, cases, read_csv
from gpu_util import announce
from oracle import sample as S
from test_sample_oracle import FIXEST_SAMPLES, reghdfe_model

pytestmark = pytest.mark.gpu


def _values(kind, n, rng):
    if kind == "small_int":  # table path
        return rng.integers(-50, 5000, n).astype(np.float64), "f64"
    if kind == "wide_float":  # sort path: the keys span most of the 64 bits
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
        v[rng.integers(0, n, n // 4)] = v[0]
        return v, "f64"
    if kind == "nan_inf":
        v = np.round(rng.standard_normal(n), 2)
        v[rng.integers(0, n, 50)] = np.nan
        v[rng.integers(0, n, 20)] = np.inf
        v[rng.integers(0, n, 20)] = -0.0
        return v, "f64"
    if kind == "i32":
        v = rng.integers(-(2 ** 31) + 1, 2 ** 31 - 1, n).astype(np.int32)
        v[rng.integers(0, n, 40)] = S.INT32_MIN
        v[rng.integers(0, n, n // 3)] = 7
        return v, "i32"
    if kind == "i64":
        v = rng.integers(-(2 ** 62), 2 ** 62, n).astype(np.int64)
        v[rng.integers(0, n, n // 2)] = 12345
        v[rng.integers(0, n, 30)] = S.INT64_MIN
        return v, "i64"
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["small_int", "wide_float", "nan_inf", "i32", "i64"])
@pytest.mark.parametrize("n", [1, 1000, 300_000])
def test_factorize_matches_the_oracle(kind, n):
    import fereg_amd.sample as FS
    rng = np.random.default_rng(n + len(kind))
    v, k = _values(kind, n, rng)
    announce("factorize", kind, n)
    codes, g, miss = FS.factorize(v)
    want, wg, wmiss = S.factorize(v, kind=k)
    np.testing.assert_array_equal(codes, want)
    assert (g, miss) == (wg, wmiss)


def test_factorize_codes_and_strict():
    import fereg_amd as F
    import fereg_amd.sample as FS
    rng = np.random.default_rng(5)
    c = rng.integers(0, 40, 10_000).astype(np.uint32)
    announce("factorize codes")
    got, g, miss = FS.factorize(FS.Codes(c, 30, base=3))
    want, wg, wm = S.factorize(c.astype(np.int64), kind="codes", base=3, nlevels=30)
    np.testing.assert_array_equal(got, want)
    assert (g, miss) == (wg, wm)
    v = np.array([np.inf, 1.0, -np.inf, 1.0])
    got, g, _ = FS.factorize(v, strict=True)
    assert list(got) == [2, 1, 0, 1] and g == 3
    with pytest.raises(F.FeregError) as e:
        FS.factorize(np.array([1.0, np.nan]), strict=True)
    assert e.value.rc == F.FH_INVALID


@pytest.mark.parametrize("gs", [(7, 13), (50_000, 40_000), (3000, 2, 900)])
def test_combine_and_restrict_match_the_oracle(gs):
    import fereg_amd.sample as FS
    rng = np.random.default_rng(len(gs))
    n = 400_000
    parts = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    announce("combine", gs)
    got, g = FS.combine(parts, gs)
    want, wg = S.combine(parts, gs)
    np.testing.assert_array_equal(got, want)
    assert g == wg
    kept = np.flatnonzero(rng.random(n) < 0.2).astype(np.uint32)
    r, gr = FS.restrict(got, g, kept)
    wr, wgr = S.restrict(want, kept)
    np.testing.assert_array_equal(r, wr)
    assert gr == wgr


@pytest.mark.parametrize("family,weighted", [("gaussian", False), ("gaussian", True), ("poisson", False),
                                             ("logit", False), ("poisson", True)])
def test_drop_singletons_matches_the_oracle(family, weighted):
    import fereg_amd.sample as FS
    rng = np.random.default_rng(11)
    n = 200_000
    gs = [60_000, 900, 40]
    terms = [S.dense_rank(rng.integers(0, g, n)) for g in gs]
    alive0 = rng.random(n) > 0.05
    y = None if family == "gaussian" else (rng.poisson(0.4, n).astype(float) if family == "poisson"
                                           else (rng.random(n) < 0.2).astype(float))
    fw = rng.integers(1, 4, n).astype(float) if weighted else None
    announce("singletons", family, weighted)
    got, s, f = FS.drop_singletons(terms, alive0, fw=fw, y_sep=y, family=family)
    want = alive0.copy()
    ws, wf = S.drop_singletons(terms, want, fw=fw, y=y, family=family)
    np.testing.assert_array_equal(got, want)
    assert (s, f) == (ws, wf)


@pytest.mark.parametrize("case,fam", FIXEST_SAMPLES, ids=lambda c: c["name"] if isinstance(c, dict) else c)
def test_fixest_samples_on_the_gpu(case, fam):
    import fereg_amd.sample as FS
    d = read_csv(case["data"])
    cols = {}
    col = lambda nm: cols.setdefault(nm, d[nm])  # one object per factor name
    announce("sample", case["name"])
    s = FS.Sample(y=d[case["y"]], x=[d[x] for x in as_list(case["x"])], fe=[tuple(col(f) for f in t["factors"]) for t in case["fe"]],
                  cluster=[tuple(col(f) for f in c) for c in case["cluster"]],
                  weights=d[case["weights"]] if case.get("weights") else None,
                  offset=d[case["offset"]] if case.get("offset") else None, family=fam, separation=fam != "gaussian")
    obs = np.array(case["expect"]["obs"], np.int64) - 1
    np.testing.assert_array_equal(s.kept(), obs)
    from test_sample_oracle import _fixest_sample
    want = _fixest_sample(case, fam)
    for q in range(s.nterms):
        c, g = s.term(q)
        np.testing.assert_array_equal(c, want["terms"][q][0])
        assert g == want["terms"][q][1]
    for q in range(s.nclusters):
        c, g = s.cluster(q)
        np.testing.assert_array_equal(c, want["clusters"][q][0])
    for k, v in want["counts"].items():
        assert s.counts[k] == v, k
    np.testing.assert_array_equal(s.gather(d[case["y"]]), d[case["y"]][obs])
    s.close()


@pytest.mark.parametrize("case", cases("reghdfe_ols.json"), ids=lambda c: c["name"])
def test_reghdfe_samples_on_the_gpu(case):
    import fereg_amd.sample as FS
    d, names, idx = reghdfe_model(case)
    w = case["weight"]
    announce("reghdfe sample", case["name"])
    s = FS.Sample(fe=[tuple(d[f] for f in t) for t in case["absorb"]], cluster=[tuple(d[f] for f in t) for t in case["cluster"]],
                  weights=d[w[0]] if w else None, weight_kind=w[1] if w else None, missing=False,
                  drop_singletons=case["singletons"]) if case["absorb"] or case["cluster"] else None
    if s is None:
        pytest.skip("no factor")
    e = case["expect"]
    keep = np.zeros(len(e["sample"]))
    keep[s.kept()] = 1.0
    np.testing.assert_array_equal(keep, np.array(e["sample"], float))
    assert s.counts["singletons"] == e["scalars"]["num_singletons"]
    s.close()


def test_sample_at_two_million_rows():
    """A sample of the configured panel's shape with missing values and singletons, against the
    oracle (codes of every term and cluster over the kept rows, counts)."""
    import fereg_amd.sample as FS
    rng = np.random.default_rng(21)
    n = 2_000_000
    indiv = rng.integers(0, 250_000, n).astype(np.float64)
    year = rng.integers(1990, 2000, n).astype(np.int64)
    firm = rng.integers(0, 30_000, n).astype(np.int32)
    firm[rng.integers(0, n, 100)] = S.INT32_MIN
    y = rng.standard_normal(n)
    y[rng.integers(0, n, 300)] = np.nan
    x = rng.standard_normal(n)
    w = rng.uniform(0.0, 2.0, n)
    w[rng.integers(0, n, 200)] = 0.0
    announce("sample 2e6")
    s = FS.Sample(y=y, x=[x], fe=[indiv, year, firm], cluster=[(indiv, year)], weights=w)
    want = S.build(n, y=y, x=[x], weights=w,
                   factors=[(indiv, "f64", 0, 0), (year, "i64", 0, 0), (firm, "i32", 0, 0)],
                   terms=[(0,), (1,), (2,)], clusters=[(0, 1)])
    np.testing.assert_array_equal(s.kept(), want["kept"])
    for q in range(3):
        c, g = s.term(q)
        np.testing.assert_array_equal(c, want["terms"][q][0])
        assert g == want["terms"][q][1]
    c, g = s.cluster(0)
    np.testing.assert_array_equal(c, want["clusters"][0][0])
    for k, v in want["counts"].items():
        assert s.counts[k] == v, k
    s.close()


def test_sample_on_device_tensors():
    import torch
    import fereg_amd.sample as FS
    rng = np.random.default_rng(4)
    n = 100_000
    a = rng.integers(0, 5000, n).astype(np.float64)
    b = rng.integers(0, 30, n).astype(np.int64)
    y = rng.standard_normal(n)
    y[:10] = np.nan
    announce("sample on device")
    ta, tb, ty = (torch.as_tensor(v, device="cuda") for v in (a, b, y))
    s = FS.Sample(y=ty, fe=[ta, tb])
    want = S.build(n, y=y, factors=[(a, "f64", 0, 0), (b, "i64", 0, 0)], terms=[(0,), (1,)])
    np.testing.assert_array_equal(s.kept(), want["kept"])
    c, g = s.term(0)
    np.testing.assert_array_equal(c, want["terms"][0][0])
    s.close()
