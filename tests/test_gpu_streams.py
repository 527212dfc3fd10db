"""The Python mirror orders the context's stream with torch's current stream: work torch
queued on a tensor before a library call happens before it, and work queued after the call
sees the call's output (GPU)."""
import numpy as np
import pytest
import torch

from oracle import cpu_ref as C
from tests.helpers import mzk

pytestmark = pytest.mark.gpu


def test_library_call_orders_after_and_before_torch_work():
    ctx = mzk.Context(0)
    try:
        assert torch.cuda.current_stream().cuda_stream != ctx.stream
        n = 1 << 22
        out = torch.zeros((n, 4), dtype=torch.int64, device="cuda")
        big = torch.ones((1 << 26,), dtype=torch.int64, device="cuda")
        for _ in range(4):        # keep torch's stream busy, then overwrite the output on it
            big.mul_(3)
        out.fill_(7)
        ctx.fr_from_seed_dev(91, 3, n, out)   # must land after fill_ (torch's stream)
        out[0, 0] = 5                          # torch work after the call sees its output
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint64)
        want = C.fr_from_seed(91, 3, n)
        assert int(got[0, 0]) == 5
        assert np.array_equal(got[1:], want[1:])
    finally:
        ctx.close()


def test_srs_and_domain_destroyed_after_their_context():
    # an SRS and a domain may outlive the context that created them (the bases are shared
    # by every context on the device): destroying them afterwards touches no context
    from oracle import bls12_381 as O
    c1, c2 = mzk.Context(0), mzk.Context(0)
    tau = C.fr_from_seed(3, 0, 1)[0]
    bases = C.srs_bases(tau, 200)
    srs = c1.srs_upload(bases)
    dom = c1.domain(3, 8)
    sc = C.fr_from_seed(4, 1, 200)
    d = torch.from_numpy(sc.view(np.int64)).cuda()
    c1.close()
    out = srs.commit_batch([(d, 200, 0)], ctx=c2)[0]     # another context still commits
    assert O.g1_jacobian_from_mont_limbs(out) == O.g1_jacobian_from_mont_limbs(C.msm_best(sc, bases))
    srs.close()
    dom.close()
    c2.close()
