"""Pins the CPU oracle: brute-force structure definitions, the explicit reduced system, the
dense dummy-variable projection and the fixest golden fixtures (no GPU)."""
import numpy as np
import pytest

from fixtures import case_inputs, cases, dense_codes, dense_residualize, dummies, fixest_ols_cases, nested_by_value
from oracle import estimate
from oracle.oracle import OracleSolver


def brute_structure(codes, gs, small_limit, fe1, bterms):
    """structure.rs:1093-1171 `check`: plain sorts define pi, cells, seg_off, codes, small_work."""
    n = len(codes[0]) if codes else 0
    c1 = codes[fe1] if fe1 >= 0 else np.zeros(n, np.uint32)
    tup = [c1] + [codes[q] for q in bterms]
    keys = [np.arange(n)] + tup[::-1]
    pi = np.lexsort(keys).astype(np.uint32)
    T = np.stack([t[pi] for t in tup], axis=1) if n else np.zeros((0, len(tup)), np.uint32)
    new = np.ones(n, bool)
    if n > 1:
        new[1:] = np.any(T[1:] != T[:-1], axis=1)
    starts = np.flatnonzero(new)
    M = len(starts)
    g1 = gs[fe1] if fe1 >= 0 else 1
    seg = np.searchsorted(T[starts, 0], np.arange(g1 + 1), side="left").astype(np.uint32)
    cc = T[starts, 1:].reshape(-1).astype(np.uint32)
    coo = None if M == n else np.append(starts, n).astype(np.uint32)
    return pi, seg, cc, coo, M


def designs():
    rng = np.random.default_rng(7)
    out = []
    n = 3000
    a = [rng.integers(0, g, n).astype(np.uint32) for g in (600, 20, 150)]
    out.append(([a[0], a[1], a[2]], [600, 20, 150], [4096, 50, 0]))
    srt = np.lexsort(a[::-1])
    out.append(([x[srt] for x in a], [600, 20, 150], [4096, 50, 0]))
    b = [rng.integers(0, g, n).astype(np.uint32) for g in (9, 40, 3, 2)]
    out.append((b, [9, 40, 3, 2], [4096, 3, 0]))
    out.append(([rng.integers(0, 2, n).astype(np.uint32), rng.integers(0, 1500, n).astype(np.uint32)], [2, 1500], [4096]))
    out.append(([rng.integers(0, 700, n).astype(np.uint32)], [700], [4096]))
    five = [rng.integers(0, g, n).astype(np.uint32) for g in (400, 3, 4, 5, 6, 2)]
    out.append((five, [400, 3, 4, 5, 6, 2], [4096, 7, 0]))
    # unused levels (codes below g but some levels empty)
    out.append(([rng.integers(0, 50, n).astype(np.uint32) * 2, rng.integers(0, 30, n).astype(np.uint32)], [101, 31], [4096, 0]))
    # elimination by assembly work (structure.rs:1266-1292)
    x = (np.arange(4000) // 40).astype(np.uint32)
    y = (np.arange(4000) % 100).astype(np.uint32)
    z = ((np.arange(4000) * 7) % 80).astype(np.uint32)
    out.append(([x, y, z], [100, 100, 80], [0, 4096]))
    return out


@pytest.mark.parametrize("di", range(8))
def test_structure_matches_bruteforce(di):
    codes, gs, limits = designs()[di]
    for lim in limits:
        s = OracleSolver(codes, gs, small_limit=lim)
        info = s.info
        term, g, off, small = s.blocks()
        st = s.structure()
        pi, seg, cc, coo, M = brute_structure(codes, gs, lim, info.fe1, list(term))
        assert info.ncells == M
        np.testing.assert_array_equal(st["pi"], pi)
        np.testing.assert_array_equal(st["seg_off"], seg)
        np.testing.assert_array_equal(st["cell_codes"], cc)
        if coo is None:
            assert st["cell_obs_off"] is None
        else:
            np.testing.assert_array_equal(st["cell_obs_off"], coo)
        # small classification and sum_seg s^2 (structure.rs:1122-1146)
        nb = len(term)
        others = [q for q in range(len(gs)) if q != info.fe1]
        tot, small_set = 0, set()
        for q in sorted(others, key=lambda q: (gs[q], q)):
            if tot + gs[q] > lim:
                break
            tot += gs[q]
            small_set.add(q)
        assert list(term) == [q for q in others if q in small_set] + [q for q in others if q not in small_set]
        work = 0
        if tot > 0:
            C = cc.reshape(M, nb)
            for sg in range(info.g1):
                d = set()
                for c in range(seg[sg], seg[sg + 1]):
                    for b in range(nb):
                        if term[b] in small_set:
                            d.add(int(off[b]) + int(C[c, b]))
                work += len(d) ** 2
        assert info.small_work == work
        demote = tot > 0 and work > 64 * M
        for b in range(nb):
            assert bool(small[b]) == (term[b] in small_set and not demote)
        for b in range(nb):
            perm, loff = s.level_cells(b)
            C = cc.reshape(M, nb)
            ref = np.lexsort([np.arange(M), C[:, b]])
            np.testing.assert_array_equal(perm, ref)
            np.testing.assert_array_equal(loff, np.searchsorted(C[ref, b], np.arange(int(g[b]) + 1)))


def test_elimination_minimizes_work():
    codes, gs, _ = designs()[7]
    n = len(codes[0])
    def e(f):
        tot = 0
        for o in range(3):
            if o == f:
                continue
            for lv in range(gs[f]):
                tot += len(set(codes[o][codes[f] == lv].tolist())) ** 2
        return tot
    best = min(range(3), key=e)
    assert best != 0
    assert OracleSolver(codes, gs).info.fe1 == best


def explicit_S(s, codes, gs, w=None):
    """S = A_RR - A_R1 N1^-1 A_1R from dummies (operator.rs:2292-2315)."""
    info = s.info
    term, g, off, small = s.blocks()
    n = len(codes[0])
    wv = np.ones(n) if w is None else w
    D1 = dummies(codes[info.fe1], gs[info.fe1])
    DR = np.concatenate([dummies(codes[q], gs[q]) for q in term], axis=1)
    A11 = D1.T @ (D1 * wv[:, None])
    n1 = np.diag(A11)
    A1R = D1.T @ (DR * wv[:, None])
    ARR = DR.T @ (DR * wv[:, None])
    n1s = np.where(n1 > 0, n1, 1.0)
    return ARR - A1R.T @ (A1R / n1s[:, None])


@pytest.mark.parametrize("weighted", [False, True])
def test_operator_explicit(weighted):
    rng = np.random.default_rng(3)
    n = 1500
    gs = [120, 9, 40]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    codes.append(codes[0] // 3)  # repeated cells via a nested term
    gs.append(40)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    s = OracleSolver(codes, gs, weights=w)
    S = explicit_S(s, codes, gs, w)
    r = 3
    p = rng.standard_normal((s.info.m, r))
    np.testing.assert_allclose(s.apply(p.reshape(-1), r).reshape(-1, r), S @ p, rtol=1e-10, atol=1e-9)
    ws = s.weight_state()
    np.testing.assert_allclose(ws["diag"], np.diag(S), rtol=1e-10, atol=1e-9)
    # rhs over rows in row order: D_R' W (v - D1 N1^-1 D1' W v)
    pi = s.structure()["pi"]
    v = rng.standard_normal((n, r))
    rows = v[pi]
    rhs, b1, wsq, wsum = s.rhs(rows.reshape(-1), r)
    wv = np.ones(n) if w is None else w
    term, *_ = s.blocks()
    D1 = dummies(codes[s.info.fe1], gs[s.info.fe1])
    DR = np.concatenate([dummies(codes[q], gs[q]) for q in term], axis=1)
    n1 = D1.T @ wv
    b1_ref = D1.T @ (v * wv[:, None])
    resid1 = v - D1 @ (b1_ref / n1[:, None])
    np.testing.assert_allclose(rhs.reshape(-1, r), DR.T @ (resid1 * wv[:, None]), rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(b1.reshape(-1, r), b1_ref, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(wsq, (wv[:, None] * v * v).sum(0), rtol=1e-12)


def test_small_factor_is_reflexive_inverse():
    """S_SS M^-1 r = r for r in the range (operator.rs:2342-2362; dense.rs:209-231)."""
    rng = np.random.default_rng(5)
    n = 2000
    gs = [300, 12, 7]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    s = OracleSolver(codes, gs)
    assert s.info.m_small == s.info.m == 19
    S = explicit_S(s, codes, gs)
    r = 2
    x = rng.standard_normal((19, r))
    rin = S @ x
    z = s.small_apply(rin.reshape(-1), r).reshape(-1, r)
    np.testing.assert_allclose(S @ z, rin, rtol=1e-9, atol=1e-9)


def residualize_designs():
    rng = np.random.default_rng(11)
    out = []
    n = 2000
    out.append(([rng.integers(0, 200, n), rng.integers(0, 30, n)], [200, 30]))
    out.append(([rng.integers(0, 200, n), rng.integers(0, 30, n), rng.integers(0, 9, n)], [200, 30, 9]))
    out.append(([rng.integers(0, 300, n), rng.integers(0, 5000, n) % 250, rng.integers(0, 6, n)], [300, 250, 6]))
    i = np.arange(n)
    out.append(([i // 10, (i % 10), (i * 7919) % 97], [200, 10, 97]))  # panel with a third term
    comp = i % 2  # two disconnected components
    out.append(([(i // 4) * 2 + 0, (rng.integers(0, 20, n) * 2 + comp)], [1000, 40]))
    out.append(([rng.integers(0, 50, n)], [50]))
    out.append(([], []))
    out.append(([i // 10, i % 10, (i // 10) // 5], [200, 10, 40]))  # nested term
    return [([np.asarray(c, np.uint32) for c in cs], gs) for cs, gs in out]


@pytest.mark.parametrize("di", range(8))
@pytest.mark.parametrize("weighted", [False, True])
def test_residualize_matches_dense_projection(di, weighted):
    codes, gs = residualize_designs()[di]
    n = len(codes[0]) if codes else 2000
    rng = np.random.default_rng(100 + di)
    w = rng.uniform(0.2, 3.0, n) if weighted else None
    s = OracleSolver(codes, gs, weights=w) if codes else OracleSolver([], [], weights=w if w is not None else np.ones(n))
    r = 3
    v = rng.standard_normal((n, r))
    v[:, 2] = 1.0 + 0.0 * v[:, 2]  # a constant column
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(v[pi])
    it, exact, coefs = s.residualize(rows, r, tol=1e-12, want_coefs=True)
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9


def test_warm_start_and_fit_rows():
    codes, gs = residualize_designs()[1]
    n = len(codes[0])
    rng = np.random.default_rng(9)
    s = OracleSolver(codes, gs)
    v = rng.standard_normal((n, 2))
    pi = s.structure()["pi"]
    rows = np.ascontiguousarray(v[pi])
    it1, _, coefs = s.residualize(rows.copy(), 2, tol=1e-12, want_coefs=True)
    rows2 = np.ascontiguousarray(v[pi])
    it2, _, _ = s.residualize(rows2, 2, tol=1e-12, warm=coefs)
    assert it2 <= 2
    np.testing.assert_allclose(rows2, dense_residualize(v, codes, gs)[pi], atol=1e-9)


def fixest_nested(case, inp):
    """fit/vcov.rs:115-136 (fixest): nested by name first, then by value unless every cluster
    variable is itself an FE term."""
    cl_names = ["^".join(c) for c in case["cluster"]]
    fe_names = ["^".join(t["factors"]) for t in case["fe"]]
    nested = [nm in cl_names or (len(t["factors"]) > 1 and any(f in cl_names for f in t["factors"]))
              for nm, t in zip(fe_names, case["fe"])]
    if not all(nested) and not all(c in fe_names for c in cl_names):
        for q in range(len(nested)):
            if not nested[q]:
                nested[q] = any(nested_by_value(inp["codes"][q], k) for k in inp["clusters"])
    return nested


def _fit_ols_oracle(case):
    inp = case_inputs(case)
    codes, gs = inp["codes"], inp["g"]
    y, X, w = inp["y"], inp["X"], inp["w"]
    n = len(y)
    block = np.concatenate([y[:, None], X], axis=1)
    if codes:
        s = OracleSolver(codes, gs, weights=w)
        pi = s.structure()["pi"]
    else:
        s = OracleSolver([], [], weights=w if w is not None else np.ones(n))
        pi = np.arange(n)
    rows = np.ascontiguousarray(block[pi])
    s.residualize(rows, rows.shape[1], tol=1e-13)
    wr = w[pi] if w is not None else None
    coef, bread, e, kept = estimate.ols_from_block(rows, wr)
    K = X.shape[1]
    np_ = estimate.nparams(len(kept), gs)
    kind = case["vcov"]
    cl = [c[pi] for c in inp["clusters"]]
    nested = fixest_nested(case, inp) if kind == "cluster" else [False] * len(codes)
    ssc = case["ssc"]
    kfix = ssc["K.fixef"]
    dfk = estimate.df_k(kind, np_, gs, nested, k_adj=ssc["K.adj"], k_fixef=kfix, n_design=len(kept))
    Xk = rows[:, 1:][:, kept]
    V = estimate.vcov(kind, bread, Xk, e, wr, n, dfk, cluster=cl if cl else None, k_adj=ssc["K.adj"],
                      g_adj=ssc["G.adj"], g_df=ssc.get("G.df", "min"))
    return coef, V, dfk, kept


@pytest.mark.parametrize("case", fixest_ols_cases(multiway=True), ids=lambda c: c["name"])
def test_fixest_ols_fixtures(case):
    coef, V, dfk, kept = _fit_ols_oracle(case)
    e = case["expect"]
    ec = np.atleast_1d(np.array(e["coef"], float))
    np.testing.assert_allclose(coef, ec, rtol=1e-8)
    np.testing.assert_allclose(V.reshape(-1), np.array(e["vcov"], float), rtol=1e-8)
    assert abs(dfk - e["df_k"]) < 1e-9
