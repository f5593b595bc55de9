"""Boundary pieces of the absorber on the GPU: term_values (solve.rs:1319-1343), fit_rows, and the
fh_ctx interrupt and log callbacks (api.rs:198-229; RfError::Interrupted = 1)."""
import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


def _design(n=50_000, seed=3, weighted=False, gs=(5000, 300, 12)):
    rng = np.random.default_rng(seed)
    gs = list(gs)
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    fe = sum(rng.standard_normal(g)[c] for c, g in zip(codes, gs))
    v = np.stack([fe + rng.standard_normal(n), rng.standard_normal(n)], 1)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    return codes, gs, v, w


@pytest.mark.parametrize("weighted", [False, True])
def test_term_values_sum_to_the_fit_and_are_normalized(weighted):
    """The term values reproduce every row's fitted FE part, and every term but the one that
    absorbs the shift has weighted mean zero."""
    import fereg_amd as F
    codes, gs, v, w = _design(weighted=weighted)
    ab = F.Absorber(codes, gs, weights=w)
    pi = ab.row_order()
    rows = np.ascontiguousarray(v[pi])
    announce("term values", weighted)
    info = ab.residualize(rows, 2, F.SolveOpts(tol=1e-12, want_coefs=True))
    fit = ab.fit_rows(info.coefs, 0)
    np.testing.assert_allclose(fit, v[pi, 0] - rows[:, 0], rtol=0, atol=1e-9)
    tv = ab.term_values(info.coefs, 0)
    assert [len(t) for t in tv] == gs
    total = sum(t[c[pi]] for t, c in zip(tv, codes))
    np.testing.assert_allclose(total, fit, rtol=0, atol=1e-9)
    wr = np.ones(len(pi)) if w is None else w[pi]
    means = [float(np.sum(wr * t[c[pi]]) / np.sum(wr)) for t, c in zip(tv, codes)]
    shift = ab.info.norm_fe1
    for q, mq in enumerate(means):
        if q != shift:
            assert abs(mq) <= 1e-10 * (1 + np.abs(tv[q]).max()), (q, means)
    ab.close()


class _Counter:
    def __init__(self, stop_after):
        self.calls, self.stop_after = 0, stop_after

    def __call__(self):
        self.calls += 1
        return self.calls > self.stop_after


@pytest.mark.parametrize("stop_after", [0, 3])
def test_interrupt_stops_a_residualization(stop_after):
    """An interrupt callback that answers true stops the call with rc 1 (before the solve, or
    between PCG iterations); the absorber then residualizes normally."""
    import fereg_amd as F
    from fereg_amd import _lib as L
    # a 5000-level block stays out of the small (dense) block: the solve iterates
    codes, gs, v, _ = _design(n=200_000, seed=4, gs=(20000, 5000, 12))
    ab = F.Absorber(codes, gs)
    pi = ab.row_order()
    rows = np.ascontiguousarray(v[pi])
    cnt = _Counter(stop_after)
    announce("interrupt", stop_after)
    with pytest.raises(F.FeregError) as e:
        ab.residualize(rows.copy(), 2, F.SolveOpts(tol=1e-14), ctx=L.CallContext(interrupt=cnt))
    assert e.value.rc == F.FH_INTERRUPTED and cnt.calls == stop_after + 1
    a1 = rows.copy()
    ab.residualize(a1, 2, F.SolveOpts(tol=1e-12))
    fresh = F.Absorber(codes, gs)
    a2 = rows.copy()
    fresh.residualize(a2, 2, F.SolveOpts(tol=1e-12))
    assert np.array_equal(a1, a2)
    ab.close()
    fresh.close()


def test_log_and_interrupt_of_a_poisson_fit():
    import fereg_amd as F
    from fereg_amd import _lib as L
    rng = np.random.default_rng(6)
    n = 40_000
    gs = [2000, 20]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    x = rng.standard_normal((n, 1))
    y = rng.poisson(np.exp(0.3 * x[:, 0])).astype(float)
    lines = []
    announce("poisson log")
    res = F.fepois(y, x, codes, gs, ctx=L.CallContext(log=lines.append, verbose=2))
    irls = [s for s in lines if "irls iteration" in s]
    assert len(irls) == res.iterations
    with pytest.raises(F.FeregError) as e:
        F.fepois(y, x, codes, gs, ctx=L.CallContext(interrupt=lambda: True))
    assert e.value.rc == F.FH_INTERRUPTED
