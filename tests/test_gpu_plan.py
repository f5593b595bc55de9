"""S*p launch plans (dense or gathered small block, lane widths) and the autotuner: every plan
computes the oracle's product; a tuned plan is cached per structure shape and can be pinned."""
import numpy as np
import pytest

from gpu_util import announce
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu


def _design(rng, n=40_000, weighted=False):
    # FE1 with 3000 levels, one small block (7 levels) and one large block (6000 levels): the
    # single small block qualifies for the dense products
    gs = [3000, 7, 6000]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    return codes, gs, w


@pytest.mark.parametrize("bucket_log2", [0, 8])
@pytest.mark.parametrize("weighted", [False, True])
def test_every_plan_matches_oracle(weighted, bucket_log2):
    # bucket_log2 = 8 cuts the 3000 FE1 levels into 12 buckets: both item shapes of the large
    # block's level pass (XCD buckets, one item per level) are exercised
    import fereg_amd as F
    rng = np.random.default_rng(11)
    codes, gs, w = _design(rng, weighted=weighted)
    o = OracleSolver(codes, gs, weights=w)
    a = F.Absorber(codes, gs, weights=w, bucket_log2=bucket_log2)
    base = a.plan(F.FH_PLAIN_FP64, 2)
    assert base["dense"] == 1 and not base["tuned"]
    nb = sum(1 for v in base["block_lanes"] if v)
    modes = (-1, 0, 1) if bucket_log2 else (-1,)
    for r in (1, 2, 3, 5):
        p = rng.standard_normal(o.info.m * r)
        yo = o.apply(p, r)
        scale = np.abs(yo).max()
        for dense in (1, 0):
            for f in (1, 2, 0.5):
                for mode in modes:
                    plan = {"dense": dense, "seg_lanes": max(1, int(base["seg_lanes"] * f)),
                            "block_lanes": [max(1, min(64, int(v * f))) if v else 0 for v in base["block_lanes"][:nb]],
                            "block_mode": [mode] * nb}
                    for prec, tol in ((F.FH_PLAIN_FP64, 1e-12), (F.FH_MIXED, 3e-5)):
                        announce("plan", weighted, bucket_log2, r, dense, f, mode, prec)
                        a.set_plan(prec, r, plan)
                        got = a.plan(prec, r)
                        assert got["tuned"] and got["dense"] == dense and got["seg_lanes"] == plan["seg_lanes"]
                        y = a.apply(p, r, precision=prec)
                        np.testing.assert_allclose(y, yo, rtol=0, atol=tol * scale)


def test_autotune_caches_and_solves():
    import fereg_amd as F
    rng = np.random.default_rng(12)
    codes, gs, w = _design(rng)
    F.lib().fh_tune_cache_clear()
    o = OracleSolver(codes, gs)
    a = F.Absorber(codes, gs, autotune=True)
    assert not a.plan(F.FH_PLAIN_FP64, 2)["tuned"]
    p = rng.standard_normal(o.info.m * 2)
    announce("autotune apply")
    y = a.apply(p, 2)
    np.testing.assert_allclose(y, o.apply(p, 2), rtol=0, atol=1e-12 * np.abs(y).max())
    tuned = a.plan(F.FH_PLAIN_FP64, 2)
    assert tuned["tuned"]
    for v in [tuned["seg_lanes"]] + [v for v in tuned["block_lanes"] if v]:
        assert v & (v - 1) == 0 and 1 <= v <= 64
    # a second absorber of the same shape takes the cached plan at its first product
    b = F.Absorber(codes, gs, autotune=True)
    b.apply(p, 2)
    assert b.plan(F.FH_PLAIN_FP64, 2) == tuned
    # residualization under the tuned plans stays at the solver's accuracy (mixed form too)
    for prec in (F.FH_PLAIN_FP64, F.FH_MIXED):
        c = F.Absorber(codes, gs, autotune=True, precision=prec)
        pi = c.row_order()
        v = rng.standard_normal((len(codes[0]), 2))
        rows = np.ascontiguousarray(v[pi])
        ro = rows.copy()
        announce("autotune residualize", prec)
        c.residualize(rows, 2, F.SolveOpts(tol=1e-10))
        o.residualize(ro, 2, tol=1e-10)
        np.testing.assert_allclose(rows, ro, atol=1e-8)
    F.lib().fh_tune_cache_clear()


def test_invalid_plans_are_rejected():
    import fereg_amd as F
    rng = np.random.default_rng(13)
    codes, gs, _ = _design(rng, n=5000)
    a = F.Absorber(codes, gs)
    with pytest.raises(Exception):
        a.set_plan(F.FH_PLAIN_FP64, 2, {"dense": 1, "seg_lanes": 3})
    with pytest.raises(Exception):
        a.set_plan(F.FH_PLAIN_FP64, 2, {"dense": 2})
    with pytest.raises(Exception):
        a.set_plan(F.FH_PLAIN_FP64, 2, {"dense": 1, "block_lanes": [128]})
    with pytest.raises(Exception):
        a.set_plan(F.FH_PLAIN_FP64, 2, {"dense": 1, "block_mode": [2]})


@pytest.mark.parametrize("n,g1", [(100_000, 100), (40_000, 3000)])
def test_dense_small_block_counts_beyond_u8(n, g1):
    """Without prior weights E holds cell counts, read from a u8 copy when every count fits; with
    100 FE1 levels, 5 small and 50 large levels most (small, large) counts exceed 255 and S*p
    must fall back to E in the solve precision (the second design stays within u8)."""
    import fereg_amd as F
    rng = np.random.default_rng(14)
    gs = [g1, 5, 50]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    o = OracleSolver(codes, gs)
    a = F.Absorber(codes, gs)
    assert a.plan(F.FH_PLAIN_FP64, 2)["dense"] == 1
    for r in (1, 2, 3):
        p = rng.standard_normal(o.info.m * r)
        yo = o.apply(p, r)
        scale = np.abs(yo).max()
        for prec, tol in ((F.FH_PLAIN_FP64, 1e-12), (F.FH_MIXED, 3e-5)):
            announce("u8 counts", n, g1, r, prec)
            np.testing.assert_allclose(a.apply(p, r, precision=prec), yo, rtol=0, atol=tol * scale)
    v = rng.standard_normal((n, 2))
    pi = a.row_order()
    rows = np.ascontiguousarray(v[pi])
    ro = rows.copy()
    a.residualize(rows, 2, F.SolveOpts(tol=1e-10))
    o.residualize(ro, 2, tol=1e-10)
    np.testing.assert_allclose(rows, ro, atol=1e-8)
