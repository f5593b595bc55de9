"""Escalation of the FE solver (solve.rs:838-891, cg.rs:254-274): a slowly converging PCG
(rho_hat > 0.6 at iteration 12) gives the first large block the exact block preconditioner and
restarts from the current iterate; the residuals stay those of the exact projection. A small-block
limit of 64 levels keeps the 1500-level ring a large (Jacobi) block until it escalates."""
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
        codes.append(rng.integers(0, extra, 2 * F))
        gs.append(extra)
    return [np.asarray(c, np.uint32) for c in codes], gs


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("weighted", [False, True])
def test_slow_ring_escalates_and_matches_the_projection(precision, weighted):
    import fereg_amd as F
    extra = None
    codes, gs = _ring(1500, extra)
    n = len(codes[0])
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    a = F.Absorber(codes, gs, weights=w, precision=precision, small_limit=64)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("escalation ring", extra, precision, weighted)
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert info.escalated, info
    ref = dense_residualize(v, codes, gs, w)[pi]
    scale = np.sqrt((v ** 2).sum(0))
    assert np.max(np.abs(rows - ref) / scale) <= 1e-9, info
    # the escalated preconditioner stays with the absorber: a second solve starts with it (and
    # does not escalate again)
    rows2 = np.ascontiguousarray(v[pi])
    info2 = a.residualize(rows2, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info2.escalated, info2
    assert info2.exact and info2.iterations == 1, info2  # the block preconditioner is S itself
    np.testing.assert_allclose(rows2, rows, rtol=0, atol=1e-9 * scale.max())


def test_two_large_blocks_do_not_escalate():
    """With a second large block the escalation is not taken: the slow ring keeps Jacobi and
    still converges to the projection."""
    import fereg_amd as F
    codes, gs = _ring(600, 400)
    n = len(codes[0])
    rng = np.random.default_rng(7)
    a = F.Absorber(codes, gs, small_limit=64)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("two large blocks")
    info = a.residualize(rows, 2, F.SolveOpts(tol=1e-12, maxiter=100_000))
    assert not info.escalated, info
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
