"""The autotuner on the device: tuning runs return OK (each candidate plan was checked
against the others inside the tuner), MSMs and NTTs after tuning equal the oracle's, and
imported plans are the ones the planners use (window bits reported by the last plan)
without changing any result (GPU)."""
import numpy as np
import pytest
import torch

from oracle import cpu_ref as C
from tests.helpers import mzk, pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.fixture(autouse=True)
def clean_table():
    mzk.tune_reset()
    yield
    mzk.tune_reset()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


def affine(jac):
    return C.g1_to_affine(np.asarray(jac, dtype=np.uint64).reshape(-1, 18))


def test_tuned_msm_equals_oracle(ctx):
    tau = plonk.mont_limbs(0x5EED)
    srs = ctx.srs_generate(tau, 1 << 14)
    mzk.tune_msm(ctx, srs, 8, 14, reps=1)
    text = mzk.tune_export()
    assert text.count("msm ") == 7
    bases = srs.download()
    for log_n in (8, 11, 14):
        n = 1 << log_n
        s = C.fr_from_seed(91, 1, n)
        got = srs.commit(dev(s), n)
        assert np.array_equal(affine(got), affine(C.msm_best(s, bases[:n])))
    srs.close()


@pytest.mark.parametrize("c_bits", [4, 8, 16])
def test_imported_window_bits_are_used(ctx, c_bits):
    n = 1 << 12
    bases = C.g1_mul_generator_many(C.fr_from_seed(5, 2, n))
    s = C.fr_from_seed(6, 1, n)
    want = affine(C.msm_best(s, bases))
    mzk.tune_import(f"msm 12 0 {c_bits} 256 0\n")
    got = ctx.msm_dev(dev(s), dev(bases), n)
    assert ctx.msm_last_plan()["c"] == c_bits
    assert np.array_equal(affine(got), want)


def test_imported_path_switch(ctx):
    tau = plonk.mont_limbs(0xABC)
    srs = ctx.srs_generate(tau, 1 << 15)
    n = 1 << 15
    s = C.fr_from_seed(8, 1, n)
    ref = affine(srs.commit(dev(s), n))
    mzk.tune_import("msm 15 0 8 1024 0\n")          # per-window buckets, c = 8
    a = srs.commit(dev(s), n)
    assert ctx.msm_last_plan()["c"] == 8
    mzk.tune_import("msm 15 1 0 0 256\n")           # shared buckets over the shifted bases
    b = srs.commit(dev(s), n)
    assert ctx.msm_last_plan()["bucket_sets"] == 1
    assert np.array_equal(affine(a), ref) and np.array_equal(affine(b), ref)
    srs.close()


def test_tuned_ntt_equals_oracle(ctx):
    mzk.tune_ntt(ctx, 12, 16, reps=1)
    assert mzk.tune_export().count("ntt ") == 5
    for log_n in (12, 13, 16):
        a = C.fr_from_seed(17, 2, 1 << log_n)
        d_in, d_out = dev(a), torch.empty((1 << log_n, 4), dtype=torch.int64, device="cuda")
        ctx.ntt(mzk.NttJob.make(d_in, d_out, log_n))
        ctx.synchronize()
        got = d_out.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, C.ntt(a, log_n))


@pytest.mark.parametrize("l1,groups", [(5, 256), (7, 4096), (9, 1024)])
def test_imported_ntt_split(ctx, l1, groups):
    log_n = 14
    a = C.fr_from_seed(18, 2, 1 << log_n)
    mzk.tune_import(f"ntt {log_n} {l1} {groups}\n")
    d_in, d_out = dev(a), torch.empty((1 << log_n, 4), dtype=torch.int64, device="cuda")
    ctx.ntt(mzk.NttJob.make(d_in, d_out, log_n))
    ctx.synchronize()
    assert np.array_equal(d_out.cpu().numpy().view(np.uint64), C.ntt(a, log_n))


@pytest.mark.parametrize("log_n,l1,groups", [(12, 4, 256), (14, 9, 4096), (16, 6, 1024)])
def test_imported_split_inverse_and_coset(ctx, log_n, l1, groups):
    """A tuned split is used by every transform of that size: inverse (lagrange_to_coeff)
    and coset extension (coeff_to_extended 2^(log_n-2) -> 2^log_n) stay equal to the oracle."""
    mzk.tune_import(f"ntt {log_n} {l1} {groups}\n")
    n = 1 << log_n
    a = C.fr_from_seed(21, 2, n)
    d_in, d_out = dev(a), torch.empty((n, 4), dtype=torch.int64, device="cuda")
    ctx.ntt(mzk.NttJob.make(d_in, d_out, log_n, mzk.NTT_INVERSE))
    ctx.synchronize()
    assert np.array_equal(d_out.cpu().numpy().view(np.uint64), C.intt(a, log_n))
    small = C.fr_from_seed(22, 2, n // 4)
    d_small = dev(small)
    ctx.ntt(mzk.NttJob.make(d_small, d_out, log_n, mzk.NTT_COSET, n_real=n // 4))
    ctx.synchronize()
    assert np.array_equal(d_out.cpu().numpy().view(np.uint64), C.coeff_to_extended(small, log_n - 2, log_n))


@pytest.mark.parametrize("n", [1, 3, 100, 4097])
def test_shared_set_channel_tree_sparse(ctx, n):
    """The shared bucket set (2^15 buckets) forced at tiny n: nearly every bucket is empty,
    so the block tree of the channel sums runs mostly over identity leaves and whole blocks
    of identities.  The commitment equals the oracle's msm_best and the tree kernels ran."""
    tau = plonk.mont_limbs(0x7EE)
    srs = ctx.srs_generate(tau, 1 << 14)   # the window-shifted tables exist from 2^14 bases on
    bases = srs.download()
    mzk.tune_import("".join(f"msm {L} 1 0 0 256\n" for L in range(15)))
    s = C.fr_from_seed(12, 1, n)
    ctx.profile(True)
    got = srs.commit(dev(s), n)
    names = set(ctx.profile_read())
    ctx.profile(False)
    assert ctx.msm_last_plan()["bucket_sets"] == 1
    assert {"k_chan_tree", "k_chan_top"} <= names
    assert np.array_equal(affine(got), affine(C.msm_best(s, bases[:n])))
    srs.close()
