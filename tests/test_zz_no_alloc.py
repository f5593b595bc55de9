"""fh_absorber_apply_noalloc (S*p with every scratch array in the library's static device array)
against fh_absorber_apply through the C ABI: the same kernels and launch plan, so the same values
bit for bit, on two calls in a row (the second on a smaller design, after the first left the
shared array dirty), with the library's device pool untouched; a product beyond the array is
refused."""
import ctypes

import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


def _panel_absorber(n, seed):
    import fereg_amd as F
    from fereg_amd import synth
    d = synth.panel(n, T=10, F=max(1, n // 100), seed=seed)
    return F.Absorber([d["indiv"], d["firm"], d["year"]], d["g"])


def _noalloc(ab, p, r, y, precision):
    from fereg_amd import _lib as L
    return L.lib().fh_absorber_apply_noalloc(ab._h, ctypes.c_void_p(p.data_ptr()), r, ctypes.c_void_p(y.data_ptr()),
                                             precision)


def test_noalloc_apply_matches_apply_on_two_calls():
    import torch
    import fereg_amd as F
    from fereg_amd import _lib as L
    big, small = _panel_absorber(2_000_000, 5), _panel_absorber(300_000, 6)
    rng = np.random.default_rng(8)
    for ab, r in ((big, 3), (small, 1)):
        m = ab.info.m
        for prec in (F.FH_PLAIN_FP64, F.FH_MIXED):
            announce("noalloc apply", ab.info.n, r, prec)
            p = rng.standard_normal(m * r)
            ref = ab.apply(p, r, precision=prec)
            p_d = torch.as_tensor(p).to("cuda:0")
            y_d = torch.full((m * r,), float("nan"), dtype=torch.float64, device="cuda:0")
            torch.cuda.synchronize()
            before = L.memory_stats()
            rc = _noalloc(ab, p_d, r, y_d, prec)
            after = L.memory_stats()
            assert rc == 0, L.lib().fh_last_error()
            assert before == after  # no block taken from or returned to the pool
            got = y_d.cpu().numpy()
            assert np.array_equal(got, ref)
    big.close()
    small.close()


def test_noalloc_apply_refuses_beyond_its_array():
    import torch
    import fereg_amd as F
    from fereg_amd import _lib as L
    ab = _panel_absorber(10_000_000, 7)  # 1e6 / 1e5 / 10 levels
    m = ab.info.m
    announce("noalloc apply limits", m)
    # FP64 with two columns fits (the documented size)
    p = torch.ones(m * 2, dtype=torch.float64, device="cuda:0")
    y = torch.empty(m * 2, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    assert _noalloc(ab, p, 2, y, F.FH_PLAIN_FP64) == 0
    # sixteen FP64 columns need 16 x 8 MB of segment sums alone: refused, nothing launched
    p = torch.ones(m * 16, dtype=torch.float64, device="cuda:0")
    y = torch.empty(m * 16, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = _noalloc(ab, p, 16, y, F.FH_PLAIN_FP64)
    assert rc == F.FH_INVALID and b"scratch" in L.lib().fh_last_error()
    ab.close()
