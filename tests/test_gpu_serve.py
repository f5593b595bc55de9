"""Serving engine: requests run concurrently on the device give the same results as alone (bitwise
when each runs as its own solve; within the tolerance when small OLS requests are packed into one
solve)."""
import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


def test_results_identical_alone_and_in_the_mix():
    import fereg_amd as F
    from fereg_amd import serve
    specs = serve.poisson_stream(16, seed=3, n_lo=2e3, n_hi=6e4, ppml_share=0.3)
    data = [serve.make_request(s) for s in specs]
    announce("serve mix", len(specs))
    srv = serve.Server(slots=6, pack=1)
    tickets = []
    for s, (codes, g, y, X) in zip(specs, data):
        tickets.append(srv.submit(s["model"], codes, g, y, X, vcov="hetero" if s["model"] == "ppml" else "iid"))
    mixed = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.rc == 0 for r in mixed)
    announce("serve alone")
    alone_srv = serve.Server(slots=1, pack=1)
    for s, (codes, g, y, X), rm in zip(specs, data, mixed):
        t = alone_srv.submit(s["model"], codes, g, y, X, vcov="hetero" if s["model"] == "ppml" else "iid")
        ra = alone_srv.wait(t)
        assert ra.rc == 0
        # fixed-order reductions: bitwise the same whatever runs beside
        np.testing.assert_array_equal(ra.coef, rm.coef)
        np.testing.assert_array_equal(ra.vcov, rm.vcov)
    alone_srv.close()
    # and the same as the direct entry points
    s, (codes, g, y, X) = specs[0], data[0]
    ab = F.Absorber([np.asarray(c, np.uint32) for c in codes], g, precision=F.FH_MIXED)
    fn = F.fepois if s["model"] == "ppml" else F.feols
    fit = fn(y, X, None, None, vcov="hetero" if s["model"] == "ppml" else "iid", opts=F.SolveOpts(tol=1e-8),
             absorber=ab)
    ab.close()
    np.testing.assert_allclose(fit.coef, mixed[0].coef, rtol=1e-12, atol=1e-14)


def test_admission_by_memory_budget():
    """A budget below one request's footprint still runs each request, alone: with four slots free
    the engine admits one request at a time, so no two service intervals overlap."""
    from fereg_amd import serve
    specs = serve.poisson_stream(4, seed=5, n_lo=2e3, n_hi=2e4, ppml_share=0.0)
    srv = serve.Server(slots=4, hbm_budget=1)
    tickets = [srv.submit(s["model"], *serve.make_request(s)) for s in specs]
    replies = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.rc == 0 for r in replies)
    spans = sorted((r.start_s, r.end_s) for r in replies)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s0 <= e0 <= s1 <= e1, spans
    # FIFO admission: the service order is the submission order
    assert [r.start_s for r in replies] == sorted(r.start_s for r in replies)


def test_admission_overlaps_with_room():
    """With room for every request, requests submitted together are in flight together."""
    from fereg_amd import serve
    specs = serve.poisson_stream(4, seed=6, n_lo=2e4, n_hi=4e4, ppml_share=0.0)
    data = [serve.make_request(s) for s in specs]
    srv = serve.Server(slots=4)
    tickets = [srv.submit(s["model"], *d) for s, d in zip(specs, data)]
    replies = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.rc == 0 for r in replies)
    first_end = min(r.end_s for r in replies)
    assert sum(r.start_s < first_end for r in replies) >= 2


def test_packed_requests_match_their_solves_alone():
    """Small OLS requests queued together are packed into one absorber and one PCG (disjoint level
    ranges, per-request column scaling); each result equals its solve alone within the tolerance,
    and the Poisson requests between them still run alone."""
    from fereg_amd import serve
    specs = serve.poisson_stream(24, seed=11, n_lo=1e3, n_hi=5e4, ppml_share=0.25)
    data = [serve.make_request(s) for s in specs]
    announce("serve packed", len(specs))
    vc = lambda s: "hetero" if s["model"] == "ppml" or s["seed"] % 2 else "iid"
    srv = serve.Server(slots=2, pack=32)
    tickets = [srv.submit(s["model"], *d, vcov=vc(s)) for s, d in zip(specs, data)]
    packed = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.rc == 0 for r in packed)
    assert max(r.packed for r in packed) > 1
    assert all(r.packed == 1 for s, r in zip(specs, packed) if s["model"] == "ppml")
    alone = serve.Server(slots=1, pack=1)
    for s, d, rp in zip(specs, data, packed):
        ra = alone.wait(alone.submit(s["model"], *d, vcov=vc(s)))
        assert ra.rc == 0 and ra.packed == 1
        np.testing.assert_array_equal(ra.kept, rp.kept)
        # relative to each matrix's scale (its near-zero off-diagonal entries carry rounding only)
        np.testing.assert_allclose(rp.coef, ra.coef, rtol=1e-7, atol=1e-9 * np.abs(ra.coef).max())
        np.testing.assert_allclose(rp.vcov, ra.vcov, rtol=1e-7, atol=1e-8 * np.abs(ra.vcov).max())
    alone.close()
