"""Escalation of the FE solver (solve.rs:685-891, cg.rs:254-274, precond/mod.rs:74-185): a slowly
converging PCG (rho_hat > 0.6 at iteration 12) replaces Jacobi by exact sparse factors -- the joint
reduced matrix over every coordinate (M = S, so the restart is exact), or each large block -- and
restarts from the current iterate; the residuals stay those of the exact projection. A small-block
limit of 64 levels keeps the rings large (Jacobi) blocks until they escalate."""
import numpy as np
import pytest

from fixtures import dense_residualize
from gpu_util import announce

pytestmark = pytest.mark.gpu


def _ring(F, extra=None, seed=0):
    # individual s is observed at firms s and s+1 (mod F): the firm graph is a ring of length F,
    # whose Laplacian converges slowly under Jacobi; a 3-level year term rides along
    rng = np.random.default_rng(seed)
    indiv = np.repeat(np.arange(F), 2)
    firm = np.stack([np.arange(F), (np.arange(F) + 1) % F], 1).reshape(-1)
    year = rng.integers(0, 3, 2 * F)
    codes = [indiv, firm, year]
    gs = [F, F, 3]
    if extra is not None:
        # a second slow ring: pairs of individuals, each observed at two adjacent levels
        codes.append(((indiv // 2) + np.tile([0, 1], F)) % extra)
        gs.append(extra)
    return [np.asarray(c, np.uint32) for c in codes], gs


def _level_sums_max(e, codes, gs, w):
    """max over terms and levels of |sum_i w_i e_i| (D'W e, zero for the projection's residual)."""
    w = np.ones(len(e)) if w is None else w
    out = 0.0
    for c, g in zip(codes, gs):
        for j in range(e.shape[1]):
            out = max(out, np.abs(np.bincount(c, w * e[:, j], minlength=g)).max())
    return out


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("weighted", [False, True])
def test_slow_ring_escalates_to_the_joint_factor(precision, weighted):
    import fereg_amd as F
    codes, gs = _ring(1500)
    n = len(codes[0])
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    a = F.Absorber(codes, gs, weights=w, precision=precision, small_limit=64)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("escalation ring joint", precision, weighted)
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info.escalated, info
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info
    # the joint factor stays with the absorber: a second solve is exact in one step
    rows2 = np.ascontiguousarray(v[pi])
    info2 = a.residualize(rows2, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info2.escalated, info2
    assert info2.exact and info2.iterations == 1, info2
    np.testing.assert_allclose(rows2, rows, rtol=0, atol=1e-9 * scale.max())


@pytest.mark.parametrize("weighted", [False, True])
def test_block_factors_without_the_joint(weighted):
    """joint off: the large block gets its own exact factor beside the dense small block
    (block-Jacobi with exact blocks); the restarted PCG converges to the projection quickly."""
    import fereg_amd as F
    codes, gs = _ring(1500)
    n = len(codes[0])
    rng = np.random.default_rng(8)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    a = F.Absorber(codes, gs, weights=w, small_limit=64, joint=False)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("escalation ring blocks", weighted)
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info.escalated, info
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info
    rows2 = np.ascontiguousarray(v[pi])
    info2 = a.residualize(rows2, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info2.escalated and not info2.exact and info2.iterations < 40, info2
    np.testing.assert_allclose(rows2, rows, rtol=0, atol=1e-9 * scale.max())


@pytest.mark.parametrize("joint", [True, False])
def test_two_large_blocks_escalate(joint):
    """A second large term next to the slow ring: the joint factor covers both (exact restart), or
    each block gets its own factor; both converge to the projection."""
    import fereg_amd as F
    codes, gs = _ring(600, 300)
    n = len(codes[0])
    rng = np.random.default_rng(7)
    a = F.Absorber(codes, gs, small_limit=64, joint=joint)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("two large blocks", joint)
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info.escalated, info
    ref = dense_residualize(v, codes, gs)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info
    rows2 = np.ascontiguousarray(v[pi])
    info2 = a.residualize(rows2, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info2.exact == joint, info2
    np.testing.assert_allclose(rows2, rows, rtol=0, atol=1e-9 * scale.max())


def test_factor_beyond_lds_and_many_columns():
    """A ring of 5000 firms: the joint factor's vector no longer fits in LDS, so the triangular
    solves run level by level in HBM; 11 columns. The residuals are checked by the projection's
    defining property, D'W e = 0 for every term."""
    import fereg_amd as F
    codes, gs = _ring(5000)
    n = len(codes[0])
    rng = np.random.default_rng(9)
    w = rng.uniform(0.5, 2.0, n)
    a = F.Absorber(codes, gs, weights=w, small_limit=64)
    v = rng.standard_normal((n, 11))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("escalation ring 5000 x 11")
    info = a.residualize(rows, 11, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info.escalated, info
    e = np.empty_like(rows)
    e[pi] = rows
    scale = np.sqrt((w[:, None] * v ** 2).sum(0)).max()
    assert _level_sums_max(e, codes, gs, w) <= 1e-9 * scale
    rows2 = np.ascontiguousarray(v[pi])
    info2 = a.residualize(rows2, 11, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info2.exact and info2.iterations == 1, info2
    np.testing.assert_allclose(rows2, rows, rtol=0, atol=1e-9 * scale)


def test_stale_factor_after_new_weights():
    """New weights keep the factor as the preconditioner (not exact for them) and the solve still
    reaches the projection for the new weights; a later escalation rebuilds it."""
    import fereg_amd as F
    codes, gs = _ring(1500)
    n = len(codes[0])
    rng = np.random.default_rng(11)
    a = F.Absorber(codes, gs, small_limit=64)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("stale factor")
    assert a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000)).escalated
    w = rng.uniform(0.5, 2.0, n)
    a.set_weights(np.ascontiguousarray(w[pi]))
    rows = np.ascontiguousarray(v[pi])
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info.exact, info
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info


def test_no_escalation_when_switched_off():
    import fereg_amd as F
    codes, gs = _ring(600)
    n = len(codes[0])
    rng = np.random.default_rng(12)
    a = F.Absorber(codes, gs, small_limit=64, sparse_chol=False)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("sparse_chol off")
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info.escalated and info.iterations > 12, info
    ref = dense_residualize(v, codes, gs)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info


def test_fast_designs_do_not_escalate():
    import fereg_amd as F
    rng = np.random.default_rng(6)
    n = 20_000
    gs = [2000, 300, 7]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    a = F.Absorber(codes, gs, small_limit=64)
    v = rng.standard_normal((n, 2))
    rows = np.ascontiguousarray(v[a.row_order()])
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-10))
    assert not info.escalated and info.iterations < 12


def test_early_probe_builds_the_joint_factor():
    """The ring's projected PCG time at the iteration-6 probe exceeds the joint factor's estimated
    cost (solve.rs:932-975), so the joint factor is built there and the restart is exact: fewer
    iterations in total than an escalation at iteration 12 could give."""
    import fereg_amd as F
    from fereg_amd._lib import CallContext
    codes, gs = _ring(1500)
    n = len(codes[0])
    rng = np.random.default_rng(13)
    a = F.Absorber(codes, gs, small_limit=64)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    lines = []
    announce("early probe")
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000),
                         ctx=CallContext(log=lines.append, verbose=2))
    print("\n".join(lines))
    assert info.escalated and info.iterations < 12, (info, lines)
    assert any("early probe" in s for s in lines), lines
    ref = dense_residualize(v, codes, gs)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info
