"""The WCR11 bootstrap oracle pinned on the CPU: SplitMix64 known answers, the Rademacher bit
order, AS241 against scipy, the fast cell-score formula against the brute-force refit (FE nested
in the clusters, where the two are the same statistic), the sign symmetry of full enumeration
and the test-inverted interval."""
import math

import numpy as np
import pytest

from fixtures import dense_residualize, dummies
from oracle import wildboot as wb


def test_splitmix64_known_answers():
    # published first outputs of SplitMix64 seeded with 0
    assert [int(x) for x in wb.splitmix64(0, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_stream_and_rademacher_bit_order():
    seed, b, G = 77, 5, 150
    # stream(seed, b): new(seed ^ b * K) with its first draw discarded (rng.rs:14-18)
    raw = wb.splitmix64(seed ^ ((b * wb.STREAM) & wb.M64), 4)
    words = wb.stream_words(seed, b, 3)
    assert [int(x) for x in words] == [int(x) for x in raw[1:4]]
    V, full = wb.rademacher(seed, b + 1, G)
    assert not full
    for c in range(G):
        bit = (int(words[c // 64]) >> (c % 64)) & 1
        assert V[b, c] == (-1.0 if bit else 1.0)
    # full enumeration: draw b, cluster c = bit c of b
    V, full = wb.rademacher(1, 5000, 12)
    assert full and V.shape == (4096, 12)
    assert V[0].tolist() == [1.0] * 12 and V[4095].tolist() == [-1.0] * 12
    assert V[5, 0] == -1.0 and V[5, 1] == 1.0 and V[5, 2] == -1.0


def test_qnorm_matches_scipy():
    from scipy.stats import norm
    for p in (1e-300, 1e-12, 1e-4, 0.025, 0.3, 0.5, 0.7, 0.975, 1 - 1e-9):
        assert abs(wb.qnorm(p) - norm.ppf(p)) <= 1e-12 * max(1.0, abs(norm.ppf(p)))


def nested_design(n, G, per, k, weighted, seed):
    """Clusters 0..G-1; FE 'store' with `per` stores inside every cluster (nested) and a
    cluster FE; k regressors."""
    rng = np.random.default_rng(seed)
    cl = rng.integers(0, G, n)
    store = cl * per + rng.integers(0, per, n)
    X = rng.standard_normal((n, k)) + 0.3 * rng.standard_normal(G)[cl][:, None]
    y = X @ np.linspace(0.2, 1.0, k) + rng.standard_normal(G * per)[store] + rng.standard_normal(n) * (1 + 0.5 * (cl % 3))
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    return y, X, cl, store, w


def fit_fwl(y, X, codes, gs, w):
    """Residualize, beta, bread and residuals by the dense projection."""
    Z = dense_residualize(np.column_stack([y, X]), codes, gs, w)
    yt, Xt = Z[:, 0], Z[:, 1:]
    wv = np.ones(len(y)) if w is None else w
    A = Xt.T @ (wv[:, None] * Xt)
    bread = np.linalg.inv(A)
    beta = bread @ (Xt.T @ (wv * yt))
    return Xt, yt - Xt @ beta, beta, bread


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("G,reps", [(40, 199), (8, 999)])
def test_fast_formula_equals_brute_force_refit(weighted, G, reps):
    n, per, k = 480, 3, 2
    y, X, cl, store, w = nested_design(n, G, per, k, weighted, seed=11 + G)
    Xt, u, beta, bread = fit_fwl(y, X, [store], [G * per], w)
    R = np.array([1.0, 0.0])
    r = 0.1
    ssc = 1.37  # any factor: the same multiplier enters both computations
    res = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=2024)
    D = dummies(store, G * per)
    stat, se, tb = wb.brute_t_boot(X, y, w, D, cl, G, R, r, reps, 2024, ssc)
    assert res["full"] == (G == 8)
    assert abs(res["stat"] - stat) <= 1e-9 * abs(stat)
    assert abs(res["se"] - se) <= 1e-9 * se
    np.testing.assert_allclose(res["t_boot"], tb, rtol=1e-8, atol=1e-9)


def test_full_enumeration_sign_symmetry_and_interval():
    G = 10
    y, X, cl, store, w = nested_design(400, G, 2, 2, False, seed=5)
    Xt, u, beta, bread = fit_fwl(y, X, [store], [G * 2], None)
    R = np.array([0.0, 1.0])
    res = wb.wcr11_result(beta, Xt, u, None, cl, G, bread, R, 0.0, 1.0, 5000, seed=1)
    tb = res["t_boot"]
    assert res["full"] and tb.size == 1 << G
    # draw b and its complement flip every sign: t is odd in V
    comp = (1 << G) - 1 - np.arange(1 << G)
    np.testing.assert_allclose(tb[comp], -tb, rtol=1e-10, atol=1e-12)
    lo, hi = res["ci"]
    assert lo < res["point"] < hi
    c = res["coefs"]

    def p_at(rv):
        d = rv - res["point"]
        st = (res["point"] - rv) / res["se"]
        return wb.compute_pvalue(st, wb.eval_affine_t(d, *c.T))[0]

    width = hi - lo
    assert p_at(res["point"]) >= 0.05
    assert p_at(lo - 0.05 * width) < 0.05 + 2.0 / tb.size and p_at(hi + 0.05 * width) < 0.05 + 2.0 / tb.size
    assert p_at(lo + 0.05 * width) >= 0.05 - 2.0 / tb.size and p_at(hi - 0.05 * width) >= 0.05 - 2.0 / tb.size


def test_pvalue_counts_with_tolerance():
    t = np.array([1.0, -2.0, 2.0 + 1e-13, 3.0, np.nan, -3.0])
    p, feas = wb.compute_pvalue(2.0, t, "two-tailed")
    assert feas == 5 and p == pytest.approx(2 / 5)
    assert wb.compute_pvalue(2.0, t, "greater")[0] == pytest.approx(1 / 5)
    assert wb.compute_pvalue(2.0, t, "less")[0] == pytest.approx(3 / 5)
    assert math.isnan(wb.compute_pvalue(float("nan"), t)[0])


def non_nested_design(n, G, T, k, weighted, seed, with_firm=True):
    """Clusters = firms (a firm FE, nested) and a year FE crossing them (not nested)."""
    rng = np.random.default_rng(seed)
    cl = rng.integers(0, G, n)
    year = rng.integers(0, T, n)
    X = rng.standard_normal((n, k)) + 0.2 * rng.standard_normal(T)[year][:, None]
    y = X @ np.linspace(0.3, 0.9, k) + rng.standard_normal(G)[cl] + rng.standard_normal(T)[year] + \
        rng.standard_normal(n) * (1 + 0.5 * (cl % 3))
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    codes = ([cl.astype(np.uint32)] if with_firm else []) + [year.astype(np.uint32)]
    gs = ([G] if with_firm else []) + [T]
    return y, X, cl, year, codes, gs, w


def path1_inputs(year, T, cl, G, w, with_firm):
    dnn = dummies(year, T)
    dt = dense_residualize(dnn, [cl.astype(np.uint32)], [G], w) if with_firm else dnn
    return dnn, dt


@pytest.mark.parametrize("with_firm", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_non_nested_projection_equals_brute_force_refit(with_firm, weighted):
    """Path 1: a year FE crossing the clusters is projected out of every bootstrap draw exactly;
    the t's equal the brute-force refits with explicit dummies."""
    G, T, k, reps = 30, 6, 2, 199
    y, X, cl, year, codes, gs, w = non_nested_design(900, G, T, k, weighted, seed=5 + weighted, with_firm=with_firm)
    Xt, u, beta, bread = fit_fwl(y, X, codes, gs, w)
    R, r, ssc = np.array([1.0, 0.0]), 0.2, 1.11
    dnn, dt = path1_inputs(year, T, cl, G, w, with_firm)
    res = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=99, dnn=dnn, dtilde=dt)
    D = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    stat, se, tb = wb.brute_t_boot(X, y, w, D, cl, G, R, r, reps, 99, ssc)
    assert abs(res["stat"] - stat) <= 1e-9 * abs(stat)
    np.testing.assert_allclose(res["t_boot"], tb, rtol=1e-8, atol=1e-10)
    # without the projection the draws differ: the projection is doing work
    plain = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=99)
    assert np.max(np.abs(plain["t_boot"] - tb)) > 1e-3


@pytest.mark.parametrize("with_firm", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_per_cluster_projection_equals_brute_force_refit(with_firm, weighted):
    """Path 2: a year term with more levels (40) than clusters (12) crossing them, projected out of
    every draw by one projection per (target, cluster); the t's equal the brute-force refits."""
    G, T, k, reps = 12, 40, 2, 199
    y, X, cl, year, codes, gs, w = non_nested_design(700, G, T, k, weighted, seed=17 + weighted, with_firm=with_firm)
    Xt, u, beta, bread = fit_fwl(y, X, codes, gs, w)
    R, r, ssc = np.array([0.0, 1.0]), 0.5, 1.07
    D = np.concatenate([dummies(c, g) for c, g in zip(codes, gs)], axis=1)
    res = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=5, p2_design=D)
    stat, se, tb = wb.brute_t_boot(X, y, w, D, cl, G, R, r, reps, 5, ssc)
    assert abs(res["stat"] - stat) <= 1e-9 * abs(stat)
    np.testing.assert_allclose(res["t_boot"], tb, rtol=1e-8, atol=1e-10)
    # Path 1 on the same design (the year dummies residualized on the nested firm term) is the
    # same projection
    dnn, dt = path1_inputs(year, T, cl, G, w, with_firm)
    p1 = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=5, dnn=dnn, dtilde=dt)
    np.testing.assert_allclose(p1["t_boot"], tb, rtol=1e-8, atol=1e-10)
    plain = wb.wcr11_result(beta, Xt, u, w, cl, G, bread, R, r, ssc, reps, seed=5)
    assert np.max(np.abs(plain["t_boot"] - tb)) > 1e-3


def _year_design(n, G, seed, weighted):
    """Clusters of stores (nested) and a year term crossing the clusters."""
    rng = np.random.default_rng(seed)
    cl = rng.integers(0, G, n)
    store = cl * 2 + rng.integers(0, 2, n)
    year = rng.integers(0, 4, n)
    X = rng.standard_normal((n, 2))
    y = X @ np.array([0.5, -0.2]) + rng.standard_normal(2 * G)[store] + rng.standard_normal(4)[year] + \
        rng.standard_normal(n) * (1.0 + 0.5 * (cl % 3))
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    return y, X, cl, store, year, w


@pytest.mark.parametrize("boot_type", [13, 31, 33])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_mnw_fast_formula_equals_brute_force(boot_type, weighted, cross):
    """The MNW algebra (ols.rs:651-1082) against refits by definition: the restricted
    leave-one-cluster-out jackknife residuals and the CRV3 variance of every bootstrap sample, with
    FE nested in the clusters and (cross) a year term crossing them through the Path-1 basis."""
    G, n = 8, 240
    y, X, cl, store, year, w = _year_design(n, G, 3 + boot_type, weighted)
    terms = [store, year] if cross else [store]
    gs = [2 * G, 4] if cross else [2 * G]
    Xt, u, beta, bread = fit_fwl(y, X, terms, gs, w)
    R = np.array([1.0, 0.0])
    r = 0.3
    ssc = 1.21
    Q = wb.mnw_basis(dummies(year, 4), dummies(store, 2 * G), w) if cross else None
    stat, se, tb = wb.mnw_fast(Xt, u, w, cl, G, bread, R, r, beta, ssc, 999, 7, boot_type, Q)
    D = np.column_stack([dummies(t, g) for t, g in zip(terms, gs)])
    bstat, bse, btb = wb.brute_mnw(X, y, w, D, cl, G, R, r, 999, 7, ssc, boot_type)
    assert abs(stat - bstat) <= 1e-9 * abs(bstat) and abs(se - bse) <= 1e-9 * bse
    np.testing.assert_allclose(tb, btb, rtol=1e-8, atol=1e-9)


def test_mnw_brute_force_with_dgp1_and_crv1_is_wcr11():
    """Boot type 11 of the brute force is the WCR11 refit oracle (the shared machinery)."""
    G = 8
    y, X, cl, store, year, w = _year_design(200, G, 4, True)
    D = dummies(store, 2 * G)
    a = wb.brute_mnw(X, y, w, D, cl, G, np.array([0.0, 1.0]), 0.0, 999, 3, 1.3, 11)
    b = wb.brute_t_boot(X, y, w, D, cl, G, np.array([0.0, 1.0]), 0.0, 999, 3, 1.3)
    assert abs(a[0] - b[0]) <= 1e-10 * abs(b[0])
    np.testing.assert_allclose(a[2], b[2], rtol=1e-10, atol=1e-12)
