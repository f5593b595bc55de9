"""GPU reduced operator, PCG and residualization against the oracle and the dense projection."""
import gc

import numpy as np
import pytest

from fixtures import dense_residualize
from gpu_util import announce
from oracle.oracle import OracleSolver
from test_oracle import residualize_designs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("weighted", [False, True])
def test_apply_matches_oracle(weighted):
    import fereg_amd as F
    rng = np.random.default_rng(3)
    n = 30_000
    gs = [3000, 9, 400]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    codes.append(codes[0] // 3)
    gs.append(1000)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    o = OracleSolver(codes, gs, weights=w)
    a = F.Absorber(codes, gs, weights=w)
    for r in (1, 2, 3, 5, 11, 16):
        p = rng.standard_normal(o.info.m * r)
        announce("apply", weighted, r)
        yo = o.apply(p, r)
        ya = a.apply(p, r)
        scale = np.abs(yo).max()
        np.testing.assert_allclose(ya, yo, rtol=0, atol=1e-12 * scale)
        y32 = a.apply(p, r, precision=F.FH_MIXED)
        np.testing.assert_allclose(y32, yo, rtol=0, atol=3e-5 * scale)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("di", range(8))
def test_residualize_matches_dense_projection(di, weighted, precision):
    import fereg_amd as F
    codes, gs = residualize_designs()[di]
    n = len(codes[0]) if codes else 2000
    rng = np.random.default_rng(100 + di)
    w = rng.uniform(0.2, 3.0, n) if weighted else None
    if not codes and w is None:
        w = np.ones(n)
    a = F.Absorber(codes, gs, weights=w, precision=precision)
    r = 3
    v = rng.standard_normal((n, r))
    v[:, 2] = 1.0
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("residualize design", di, "weighted", weighted, "precision", precision)
    info = a.residualize(rows, r, F.SolveOpts(tol=1e-12))
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info


def test_matches_oracle_iterations_and_coefs():
    import fereg_amd as F
    codes, gs = residualize_designs()[1]
    n = len(codes[0])
    rng = np.random.default_rng(4)
    v = rng.standard_normal((n, 2))
    o = OracleSolver(codes, gs)
    a = F.Absorber(codes, gs)
    pi = a.row_order()
    ro = np.ascontiguousarray(v[pi])
    ra = ro.copy()
    it_o, _, co = o.residualize(ro, 2, tol=1e-10, want_coefs=True)
    announce("oracle iterations")
    info = a.residualize(ra, 2, F.SolveOpts(tol=1e-10, want_coefs=True))
    np.testing.assert_allclose(ra, ro, atol=1e-9)
    # same algorithm and rules: the iteration counts agree up to rounding of the stopping test
    assert abs(info.iterations - it_o) <= 1
    # the fit per row (unique, unlike the FE coefficients) agrees
    fa = a.fit_rows(info.coefs, 0)
    resid_from_fit = v[pi, 0] - fa
    np.testing.assert_allclose(resid_from_fit, ra[:, 0], atol=1e-9)


def test_warm_start_and_weights_refresh():
    import fereg_amd as F
    codes, gs = residualize_designs()[2]
    n = len(codes[0])
    rng = np.random.default_rng(12)
    a = F.Absorber(codes, gs)
    pi = a.row_order()
    v = rng.standard_normal((n, 2))
    prev = None
    for rnd in range(3):
        w = rng.uniform(0.5, 2.0, n)
        a.set_weights(w[pi])
        rows = np.ascontiguousarray(v[pi])
        announce("warm round", rnd)
        info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, want_coefs=True), warm=prev)
        prev = info.coefs
        np.testing.assert_allclose(rows, dense_residualize(v, codes, gs, w)[pi], atol=1e-9)


def test_many_columns_are_chunked():
    import fereg_amd as F
    codes, gs = residualize_designs()[1]
    n = len(codes[0])
    rng = np.random.default_rng(2)
    a = F.Absorber(codes, gs)
    r = 21
    v = rng.standard_normal((n, r))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("21 columns")
    a.residualize(rows, r, F.SolveOpts(tol=1e-12))
    np.testing.assert_allclose(rows, dense_residualize(v, codes, gs)[pi], atol=1e-9)


def test_maxiter_is_enforced():
    """tests/solver.rs:489-500: maxiter exceeded -> Convergence (430)."""
    import fereg_amd as F
    n = 10000
    i = np.arange(n)
    # ring of 5000 firm levels (beyond the dense small limit): slow Jacobi convergence
    a_code = (i // 2).astype(np.uint32)
    b_code = (((i // 2) + (i % 2)) % 5000).astype(np.uint32)
    c_code = (i % 7).astype(np.uint32)
    a = F.Absorber([a_code, b_code, c_code], [5000, 5000, 7])
    assert a.info.m_small == 7
    rows = np.ascontiguousarray(np.random.default_rng(0).standard_normal((n, 1))[a.row_order()])
    announce("maxiter")
    with pytest.raises(F.FeregError) as e:
        a.residualize(rows, 1, F.SolveOpts(tol=1e-14, maxiter=5))
    assert e.value.rc == 430


def test_device_resident_block():
    import torch
    import fereg_amd as F
    codes, gs = residualize_designs()[1]
    n = len(codes[0])
    rng = np.random.default_rng(21)
    v = rng.standard_normal((n, 2))
    dcodes = [torch.as_tensor(c.astype(np.int32)).cuda() for c in codes]
    a = F.Absorber(dcodes, gs)
    pi = a.row_order()
    z = torch.as_tensor(np.ascontiguousarray(v[pi])).cuda()
    announce("device block")
    a.residualize(z, 2, F.SolveOpts(tol=1e-12))
    torch.cuda.synchronize()
    np.testing.assert_allclose(z.cpu().numpy(), dense_residualize(v, codes, gs)[pi], atol=1e-9)


@pytest.mark.parametrize("g1,bucket_log2", [(300_000, 17), (600_000, -1)])
@pytest.mark.parametrize("weighted", [False, True])
def test_bucketed_level_lists_match_oracle(weighted, g1, bucket_log2):
    """FE1 with more than 2^17 levels: the level passes run over several segment buckets (2^17
    configured: 3 buckets; automatic from 2^19 levels: one per XCD, 75000 levels each)."""
    import fereg_amd as F
    rng = np.random.default_rng(77)
    n = 2 * g1
    i = np.arange(n)
    codes = [(i // 2).astype(np.uint32), rng.integers(0, 5000, n).astype(np.uint32), (i % 2).astype(np.uint32),
             rng.integers(0, 3, n).astype(np.uint32)]
    gs = [g1, 5000, 2, 3]
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    o = OracleSolver(codes, gs, weights=w)
    a = F.Absorber(codes, gs, weights=w, bucket_log2=bucket_log2)
    assert a.info.g1 == g1 and a.info.m_small == 5
    for r in (1, 3):
        p = rng.standard_normal(o.info.m * r)
        announce("bucketed apply", weighted, r)
        yo = o.apply(p, r)
        np.testing.assert_allclose(a.apply(p, r), yo, rtol=0, atol=1e-12 * np.abs(yo).max())
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    ro = np.ascontiguousarray(v[pi])
    ra = ro.copy()
    o.residualize(ro, 2, tol=1e-11)
    announce("bucketed residualize", weighted)
    a.residualize(ra, 2, F.SolveOpts(tol=1e-11))
    np.testing.assert_allclose(ra, ro, atol=1e-8 * np.abs(ro).max())


def test_pool_reuses_and_releases_scratch():
    """Scratch goes back to the stream-ordered pool when an absorber is closed, is reused by the
    next absorber (no growth over repeated fits) and is returned to the device on request."""
    import fereg_amd as F
    codes, gs = residualize_designs()[2]
    n = len(codes[0])
    rng = np.random.default_rng(5)
    v = rng.standard_normal((n, 2))
    ref = dense_residualize(v, codes, gs)
    # absorbers of earlier tests that only the cyclic collector frees would otherwise return
    # their memory in the middle of the loop
    gc.collect()
    base_use, _ = F.memory_stats()
    peaks = []
    for rep in range(3):
        a = F.Absorber(codes, gs, precision=F.FH_MIXED)
        pi = a.row_order()
        z = np.ascontiguousarray(v[pi])
        announce("pool rep", rep)
        a.residualize(z, 2, F.SolveOpts(tol=1e-12))
        np.testing.assert_allclose(z, ref[pi], atol=1e-9)
        use, cached = F.memory_stats()
        peaks.append(use + cached)
        a.close()
        assert F.memory_stats()[0] == base_use
    assert peaks[2] == peaks[1]
    F.release_cached_memory()
    assert F.memory_stats()[1] == 0


@pytest.mark.parametrize("design,weighted", [("simple", False), ("difficult", False), ("simple", True)])
def test_mixed_form_meets_the_fp64_accuracy(design, weighted):
    """At tol 1e-8 the mixed form's last correction is usually accepted from its FP32 recursive
    residual (a rounding bound instead of an FP64 S*p): its residualized columns must stay as
    close to a tight FP64 solve as the FP64 form at the same tolerance."""
    import fereg_amd as F
    from fereg_amd import synth
    d = synth.panel(200_000, T=10, F=2_000, design=design, seed=31)
    codes = [d["indiv"], d["firm"], d["year"]]
    n = d["n"]
    rng = np.random.default_rng(32)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    v = np.column_stack([d["y"], d["X"][:, 0], rng.standard_normal(n)])
    out = {}
    for name, prec, tol in (("ref", F.FH_PLAIN_FP64, 1e-14), ("fp64", F.FH_PLAIN_FP64, 1e-8),
                            ("mixed", F.FH_MIXED, 1e-8)):
        a = F.Absorber(codes, d["g"], weights=w, precision=prec)
        pi = a.row_order()
        rows = np.ascontiguousarray(v[pi])
        announce("mixed accuracy", design, weighted, name)
        a.residualize(rows, 3, F.SolveOpts(tol=tol, maxiter=10_000))
        out[name] = rows
        a.close()
    scale = np.sqrt((out["ref"] ** 2).sum(0))
    err = {k: np.max(np.abs(out[k] - out["ref"]) / scale) for k in ("fp64", "mixed")}
    assert err["mixed"] <= max(10 * err["fp64"], 1e-10), err
    assert err["mixed"] <= 1e-7, err


@pytest.mark.parametrize("precision", [0, 1])
def test_empty_levels_with_the_dense_small_block(precision):
    """Codes need not use every level: FE1 levels without rows (their segments are empty, so
    n1 = 0), unused large-block levels and an unused small level, with one small block of 12
    levels (the dense small-row route). S*p and the residuals match the oracle, and the
    coefficients of the empty FE1 levels are 0."""
    import fereg_amd as F
    rng = np.random.default_rng(44)
    n = 200_000
    gs = [20000, 5000, 12]
    codes = [rng.integers(0, 19000, n).astype(np.uint32), rng.integers(0, 4900, n).astype(np.uint32),
             rng.integers(0, 11, n).astype(np.uint32)]
    o = OracleSolver(codes, gs)
    a = F.Absorber(codes, gs, precision=precision)
    m = o.info.m
    for r in (1, 2):
        p = rng.standard_normal(m * r)
        announce("empty levels apply", precision, r)
        yo = o.apply(p, r)
        ya = a.apply(p, r, precision=precision)
        scale = np.abs(yo).max()
        assert np.all(np.isfinite(ya))
        np.testing.assert_allclose(ya, yo, rtol=0, atol=(1e-12 if precision == 0 else 3e-5) * scale)
    v = rng.standard_normal((n, 2)) + rng.standard_normal(20000)[codes[0]][:, None]
    pi = a.row_order()
    ro = np.ascontiguousarray(v[pi])
    rg = ro.copy()
    o.residualize(ro, 2, tol=1e-12)
    announce("empty levels residualize", precision)
    info = a.residualize(rg, 2, F.SolveOpts(tol=1e-12, want_coefs=True))
    np.testing.assert_allclose(rg, ro, rtol=0, atol=1e-9 * np.abs(ro).max())
    alpha = info.coefs.data.reshape(-1, 2)[m:]
    fe1 = int(a.info.fe1)
    empty = np.setdiff1d(np.arange(gs[fe1]), codes[fe1])
    assert len(empty) > 0 and np.all(alpha[empty] == 0.0) and np.all(np.isfinite(alpha))
    a.close()


@pytest.mark.parametrize("di", [2, 3])
@pytest.mark.parametrize("dense_limit", [-1, 4096])
def test_small_blocks_dense_or_jacobi(di, dense_limit):
    """Small blocks of more than 64 levels in all (256 and 107 here): by default they keep the
    structure's order but are preconditioned like large blocks; with a higher dense limit their
    super-block is inverted densely on the GPU (pivot sweeps). Both reach the projection."""
    import fereg_amd as F
    codes, gs = residualize_designs()[di]
    n = len(codes[0])
    rng = np.random.default_rng(200 + di)
    w = rng.uniform(0.2, 3.0, n)
    a = F.Absorber(codes, gs, weights=w, dense_limit=dense_limit)
    assert a.info.m_small == sum(g for g in gs[1:])  # the structure's split is the reference's
    assert a.info.dense_small == (0 if dense_limit < 0 else 1)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("dense limit", di, dense_limit)
    a.residualize(rows, 2, F.SolveOpts(tol=1e-12))
    ref = dense_residualize(v, codes, gs, w)[pi]
    np.testing.assert_allclose(rows, ref, atol=1e-9)
