"""Device inputs handed to the library are read on the absorber's own stream, which does not wait
for torch's: the mirror drains torch's current stream first. A GPU delay queued on torch's stream
ahead of the copy that fills the input makes an unsynchronized read see the stale contents."""
import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


def _absorber_and_block(torch, F, n=200_000, r=3, seed=3):
    rng = np.random.default_rng(seed)
    gs = [20_000, 500, 7]
    dev = torch.device("cuda:0")
    # every level observed (factor-encoded codes, as the reference's input layer hands over)
    codes = [torch.as_tensor(rng.permutation(np.arange(n) % g).astype(np.int32)).to(dev) for g in gs]
    ab = F.Absorber(codes, gs)
    base = torch.as_tensor(rng.standard_normal((n, r))).to(dev)
    torch.cuda.synchronize()
    return ab, base


def test_residualize_waits_for_the_producing_stream():
    import torch
    import fereg_amd as F
    ab, base = _absorber_and_block(torch, F)
    opts = F.SolveOpts(tol=1e-12, want_coefs=True)
    ref = base.clone()
    torch.cuda.synchronize()
    info_ref = ab.residualize(ref, 3, opts)
    announce("residualize after a queued copy")
    z = torch.zeros_like(base)
    torch.cuda.synchronize()
    torch.cuda._sleep(20_000_000)  # bounded spin on torch's stream, ahead of the copy
    z.copy_(base)
    info = ab.residualize(z, 3, opts)  # the coefficient output is also allocated on torch's stream
    assert float((z - ref).abs().max()) <= 1e-12
    assert float((info.coefs.data - info_ref.coefs.data).abs().max()) <= 1e-12
    ab.close()


def test_fit_waits_for_the_producing_stream():
    import torch
    import fereg_amd as F
    ab, base = _absorber_and_block(torch, F, r=2, seed=4)
    y0, X0 = base[:, 0].contiguous(), base[:, 1:].contiguous()
    ref = F.feols(y0, X0, None, None, absorber=ab)
    announce("feols after a queued copy")
    y, X = torch.zeros_like(y0), torch.zeros_like(X0)
    torch.cuda.synchronize()
    torch.cuda._sleep(20_000_000)
    y.copy_(y0)
    X.copy_(X0)
    res = F.feols(y, X, None, None, absorber=ab)
    np.testing.assert_allclose(res.coef, ref.coef, rtol=1e-13, atol=0)
    ab.close()
