"""The device against the committed fixtures (tests/golden/fixtures.json): forward NTTs of
the seeded stream at 2^16, 2^20, 2^22 reproduce the output digests, and the benchmark's
SRS commitment at 2^16 and 2^20 reproduces the closed-form point (GPU)."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

from oracle import bls12_381 as O
from tests.helpers import mzk

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


@pytest.fixture(scope="module")
def ctx():
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("entry", FIX["ntt"], ids=lambda e: f"2^{e['log_n']}")
def test_device_ntt_digest(ctx, entry):
    log_n = entry["log_n"]
    n = 1 << log_n
    a = torch.zeros((n, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(entry["seed"], entry["stream"], n, a)
    out = torch.empty_like(a)
    ctx.ntt(mzk.NttJob.make(a, out, log_n))
    ctx.synchronize()
    y = out.cpu().numpy().view(np.uint64)
    assert hashlib.sha256(y.astype("<u8").tobytes()).hexdigest() == entry["sha256"]


@pytest.mark.parametrize("entry", FIX["msm"], ids=lambda e: f"2^{e['log_n']}")
def test_device_srs_commit(ctx, entry):
    n = 1 << entry["log_n"]
    tau = torch.zeros((1, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(entry["seed"], entry["tau_stream"], 1, tau)
    ctx.synchronize()
    srs = ctx.srs_generate(tau.cpu().numpy().view(np.uint64)[0].copy(), n)
    sc = torch.zeros((n, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(entry["seed"], entry["scalar_stream"], n, sc)
    ctx.synchronize()
    jac = srs.commit(sc, n)
    P = O.g1_affine_from_mont_limbs(mzk.g1_to_affine(jac))
    assert O.g1_to_compressed(P).hex() == entry["compressed"]
    srs.close()
