"""Batched commitments and batched NTTs against per-call results and the oracle (GPU)."""
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import empty_dev, from_dev, mzk, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def aff(j):
    return O.g1_jacobian_from_mont_limbs(j)


@pytest.mark.parametrize("shifted", [True, False])
def test_msm_batch_matches_per_call(ctx, shifted):
    tau = C.fr_from_seed(3, 0, 1)[0]
    n_srs = 1 << 17 if shifted else 12000       # below 2^14 no shifted copies are built
    bases = C.srs_bases(tau, n_srs)
    srs = ctx.srs_upload(bases)
    rng = random.Random(7)
    sizes = [int(2 ** rng.uniform(4, 16 if shifted else 13)) for _ in range(10)] + [0, 5, 1]
    jobs, host = [], []
    for i, n in enumerate(sizes):
        off = rng.randrange(0, n_srs - n + 1)
        sc = C.fr_from_seed(100 + i, 1, max(n, 1))[:n]
        if i == 11:
            sc[:] = 0                               # all-zero scalars: identity
        host.append((sc, off))
        jobs.append((to_dev(sc) if n else None, n, off))
    outs = srs.commit_batch(jobs)
    for (sc, off), (d, n, _), out in zip(host, jobs, outs):
        expect = aff(C.msm_best(sc, bases[off:off + n])) if n and sc.any() else None
        assert aff(out) == expect
        if n:
            assert aff(srs.commit(d, n, offset=off)) == expect
    srs.close()


@pytest.mark.parametrize("nj", [2, 3, 4, 5])
def test_few_shared_sets_against_oracle(ctx, nj):
    """Two to four shared bucket sets are reduced by the block tree, a set per block range
    and its channels by set; five by segments of four buckets. Each commit against the C
    restatement of msm_best and against its single commit."""
    tau = C.fr_from_seed(4, 0, 1)[0]
    n_srs = 1 << 16
    bases = C.srs_bases(tau, n_srs)
    srs = ctx.srs_upload(bases)
    rng = random.Random(nj)
    jobs, host = [], []
    for i in range(nj):
        n = rng.randrange(1 << 15, (1 << 15) + 6000)   # shared sets once jobs average 2^15 terms
        off = rng.randrange(0, n_srs - n + 1)
        sc = C.fr_from_seed(200 + 10 * nj + i, 1, n)
        host.append((sc, off))
        jobs.append((to_dev(sc), n, off))
    outs = srs.commit_batch(jobs)
    assert ctx.msm_last_plan()["bucket_sets"] == nj
    for (sc, off), (d, n, _), out in zip(host, jobs, outs):
        expect = aff(C.msm_best(sc, bases[off:off + n]))
        assert aff(out) == expect, (nj, n, off)
        assert aff(srs.commit(d, n, offset=off)) == expect
    srs.close()


def test_msm_batch_rejects_out_of_range_job(ctx):
    tau = C.fr_from_seed(3, 0, 1)[0]
    srs = ctx.srs_upload(C.srs_bases(tau, 100))
    d = to_dev(C.fr_from_seed(1, 1, 50))
    with pytest.raises(mzk.MzkError):
        srs.commit_batch([(d, 50, 0), (d, 50, 60)])
    srs.close()


def _oracle_transform(vals, log_n, flags, n_real):
    inverse = flags & mzk.NTT_INVERSE
    coset = flags & mzk.NTT_COSET
    if inverse and coset:
        return C.extended_to_coeff(vals, log_n)
    if inverse:
        return C.intt(vals, log_n)
    if coset:
        if n_real & (n_real - 1) == 0:
            k = n_real.bit_length() - 1
            return C.coeff_to_extended(vals[:n_real], k, log_n)
        ints = O.distribute_powers_zeta(C.fr_rows_to_ints(vals[:n_real]), True)
        padded = C.ints_to_fr_rows(ints + [0] * ((1 << log_n) - n_real))
        return C.ntt(padded, log_n)
    padded = np.zeros((1 << log_n, 4), dtype=np.uint64)
    padded[:n_real] = vals[:n_real]
    return C.ntt(padded, log_n)


def test_ntt_batch_mixed_sizes_and_flags(ctx):
    specs = [(3, 0, 8), (5, mzk.NTT_COSET, 8), (7, mzk.NTT_COSET, 100), (11, mzk.NTT_INVERSE, 2048),
             (12, mzk.NTT_INVERSE | mzk.NTT_COSET, 4096), (13, 0, 5000), (16, mzk.NTT_COSET, 1 << 14),
             (18, mzk.NTT_INVERSE, 1 << 18), (20, mzk.NTT_COSET, 1 << 18), (1, 0, 2)]
    ins, outs, jobs, expect = [], [], [], []
    for i, (log_n, flags, n_real) in enumerate(specs):
        vals = C.fr_from_seed(200 + i, log_n, n_real)
        din = to_dev(vals)
        dout = empty_dev(1 << log_n, 4)
        ins.append(din)
        outs.append(dout)
        jobs.append(mzk.NttJob.make(din, dout, log_n, flags, n_real))
        expect.append(_oracle_transform(vals, log_n, flags, n_real))
    ctx.ntt_batch(jobs)
    ctx.synchronize()
    for o, e in zip(outs, expect):
        assert np.array_equal(from_dev(o), e)
    # the per-call form on the same jobs (fresh outputs) agrees
    for j, (log_n, _, _), e in zip(jobs, specs, expect):
        dout = empty_dev(1 << log_n, 4)
        j.d_out = dout.data_ptr()
        ctx.ntt(j)
        ctx.synchronize()
        assert np.array_equal(from_dev(dout), e)


def test_ntt_batch_large_launch(ctx):
    # a launch of >= 2^21 points runs its stages two per LDS round trip: even and odd
    # sub-DFT sizes (2^20 = 2^10 x 2^10, 2^21 = 2^11 x 2^10), forward, coset, inverse
    specs = [(20, 0, 1 << 20), (20, mzk.NTT_COSET, 1 << 18), (21, mzk.NTT_INVERSE, 1 << 21)]
    outs, jobs, expect = [], [], []
    for i, (log_n, flags, n_real) in enumerate(specs):
        vals = C.fr_from_seed(300 + i, log_n, n_real)
        din = to_dev(vals)
        dout = empty_dev(1 << log_n, 4)
        outs.append((din, dout))
        jobs.append(mzk.NttJob.make(din, dout, log_n, flags, n_real))
        expect.append(_oracle_transform(vals, log_n, flags, n_real))
    print("\n[ntt batch 2^20, 2^20 coset, 2^21 inverse]", flush=True)
    ctx.ntt_batch(jobs)
    ctx.synchronize()
    for (_, o), e in zip(outs, expect):
        assert np.array_equal(from_dev(o), e)


def test_ntt_in_place_two_pass(ctx):
    vals = C.fr_from_seed(9, 9, 1 << 15)
    d = to_dev(vals)
    ctx.ntt(mzk.NttJob.make(d, d, 15, mzk.NTT_COSET))
    ctx.synchronize()
    assert np.array_equal(from_dev(d), C.coeff_to_extended(vals, 15, 15))


def test_ntt_job_validation(ctx):
    a = empty_dev(16, 4)
    with pytest.raises(mzk.MzkError):
        ctx.ntt(mzk.NttJob.make(a, a, 4, mzk.NTT_INVERSE, 8))       # inverse needs 2^log_n inputs
    with pytest.raises(mzk.MzkError):
        ctx.ntt(mzk.NttJob.make(a, a, 33, 0, 16))                    # beyond 2^32 (S = 32)
    with pytest.raises(mzk.MzkError):
        ctx.ntt(mzk.NttJob.make(a, a, 4, 8, 16))                     # unknown flag
