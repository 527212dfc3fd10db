"""G1 MSM on the device against the oracle (GPU).

Oracles: the C restatement of msm_best / msm_parallel (itself pinned in
test_oracle_c.py), and for SRS bases the closed form [sum_i s_i tau^i] G, which
pins the full 2^20 metric configuration independently of any MSM code.
"""
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import from_dev, mzk, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def srs16():
    tau = C.fr_from_seed(42, 0, 1)[0]
    return tau, C.srs_bases(tau, 1 << 16)


def aff(j):
    return O.g1_jacobian_from_mont_limbs(j)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 31, 100, 1000, 4096, 8200, 1 << 16])
def test_msm_matches_msm_best(ctx, srs16, n):
    _, bases = srs16
    sc = C.fr_from_seed(7, n, n)
    bs = bases[:n].copy()
    if n > 10:
        sc[3] = 0          # zero scalar (msm_specific filters it)
        bs[4] = 0          # identity base (msm_best drops it)
    expect = aff(C.msm_best(sc, bs)) if n else None
    assert aff(ctx.msm_dev(to_dev(sc), to_dev(bs), n)) == expect if n else True
    assert aff(ctx.msm_best(sc, bs)) == expect
    assert aff(ctx.msm_specific(sc, bs)) == expect


def test_msm_exceptional_buckets(ctx, srs16):
    # three distinct points and their negatives with small scalars: buckets see
    # repeated points (doubling) and cancellations (P + (-P))
    _, bases = srs16
    rng = random.Random(3)
    n = 8200
    distinct = bases[:3]
    neg = np.array([O.g1_affine_mont_limbs(O.g1_neg(O.g1_affine_from_mont_limbs(r))) for r in distinct],
                   dtype=np.uint64)
    choice = [rng.randrange(6) for _ in range(n)]
    bs = np.array([distinct[c] if c < 3 else neg[c - 3] for c in choice], dtype=np.uint64)
    sc = C.ints_to_fr_rows([rng.randrange(1 << 20) for _ in range(n)])
    assert aff(ctx.msm_best(sc, bs)) == aff(C.msm_best(sc, bs))


def test_msm_all_zero_scalars_is_identity(ctx, srs16):
    _, bases = srs16
    sc = np.zeros((100, 4), dtype=np.uint64)
    out = ctx.msm_specific(sc, bases[:100])
    assert aff(out) is None
    assert int(out[12:].sum()) == 0


def test_msm_extreme_scalars(ctx, srs16):
    _, bases = srs16
    vals = [O.R_MOD - 1, 1, 2, (1 << 254), (O.R_MOD - 1) // 2, 0, 3]
    sc = C.ints_to_fr_rows(vals)
    bs = bases[:len(vals)].copy()
    assert aff(ctx.msm_best(sc, bs)) == aff(C.msm_parallel(sc, bs))


def test_msm_length_mismatch_raises(ctx, srs16):
    _, bases = srs16
    with pytest.raises(mzk.MzkError):
        ctx.msm_best(np.zeros((3, 4), dtype=np.uint64), bases[:4])


def test_srs_generation_matches_oracle(ctx, srs16):
    tau, bases = srs16
    n = 5000
    srs = ctx.srs_generate(tau, n)
    got = srs.download()
    assert np.array_equal(got, bases[:n])
    # spot-check against the Python oracle's textbook scalar multiplication
    t = O.fr_from_mont(O.from_limbs(int(v) for v in tau))
    for i in [0, 1, 4999]:
        assert O.g1_affine_from_mont_limbs(got[i]) == O.srs_point(t, i)
    srs.close()


def test_resident_srs_commit_with_offset(ctx, srs16):
    tau, bases = srs16
    srs = ctx.srs_upload(bases[:20000])
    sc = C.fr_from_seed(11, 1, 9000)
    d = to_dev(sc)
    assert aff(srs.commit(d, 9000, offset=0)) == aff(C.msm_best(sc, bases[:9000]))
    assert aff(srs.commit(d, 9000, offset=777)) == aff(C.msm_best(sc, bases[777:9777]))
    with pytest.raises(mzk.MzkError):
        srs.commit(d, 9000, offset=15000)
    srs.close()


def test_resident_srs_shared_bucket_path_with_offset(ctx, srs16):
    """Commits of >= 2^14 terms use the window-shifted bases and one shared bucket
    set; check it (with an offset into the SRS) against msm_best."""
    tau, bases = srs16
    srs = ctx.srs_upload(bases[:40000])
    n = 20000
    sc = C.fr_from_seed(12, 1, n)
    sc[5] = 0
    d = to_dev(sc)
    for off in (0, 777, 20000):
        assert aff(srs.commit(d, n, offset=off)) == aff(C.msm_best(sc, bases[off:off + n]))
    assert ctx.msm_last_plan()["bucket_sets"] == 1
    srs.close()


def test_shared_bucket_path_edge_scalars(ctx, srs16):
    """Edge scalars through the shared bucket set (every window's digits of one scalar
    in one set): values on both sides of (r - 1) / 2, r - 1, r - 2, 0, 1, 2, 2^253,
    2^254 and window-boundary values scattered among random scalars, the same scalars
    negated, and all scalars in the upper half, against msm_best."""
    tau, bases = srs16
    srs = ctx.srs_upload(bases[:40000])
    n = 20000
    sc = C.fr_from_seed(13, 1, n)
    R = O.R_MOD
    half = (R - 1) // 2
    edge = [0, 1, 2, half - 1, half, half + 1, half + 2, R - 1, R - 2, 1 << 253, 1 << 254,
            (1 << 254) - 1, R - (1 << 253), (1 << 17) - 1, 1 << 16, R - (1 << 16)]
    rows = C.ints_to_fr_rows(edge)
    for k, row in enumerate(rows):
        sc[k * 1237 % n] = row
        sc[n - 1 - k] = row
    d = to_dev(sc)
    assert aff(srs.commit(d, n)) == aff(C.msm_best(sc, bases[:n]))
    # the same scalars negated: each term's point negated, the sum too
    neg = C.ints_to_fr_rows([(R - v) % R for v in C.fr_rows_to_ints(sc)])
    assert aff(srs.commit(to_dev(neg), n)) == aff(C.msm_best(neg, bases[:n]))
    # all scalars in the upper half at once
    up = C.ints_to_fr_rows([half + 1 + (v % half) for v in C.fr_rows_to_ints(sc)])
    assert aff(srs.commit(to_dev(up), n)) == aff(C.msm_best(up, bases[:n]))
    assert ctx.msm_last_plan()["bucket_sets"] == 1
    srs.close()


def test_msm_2_20_closed_form(ctx):
    """The metric configuration: n = 2^20, bases [tau^i]G, checked against
    [poly_s(tau)] G (one Horner pass and one scalar multiplication)."""
    n = 1 << 20
    tau = C.fr_from_seed(2024, 0, 1)[0]
    srs = ctx.srs_generate(tau, n)
    sc_dev = to_dev(np.zeros((n, 4), dtype=np.uint64))
    ctx.fr_from_seed_dev(2024, 1, n, sc_dev)
    ctx.synchronize()
    sc = from_dev(sc_dev)
    assert np.array_equal(sc[:64], C.fr_from_seed(2024, 1, 64))
    expect = O.g1_affine_from_mont_limbs(C.g1_mul_generator(C.poly_eval(sc, tau)))
    assert aff(srs.commit(sc_dev, n)) == expect
    # spot-check the generated bases themselves
    sample = srs.download(n - 3, 3)
    t = O.fr_from_mont(O.from_limbs(int(v) for v in tau))
    assert O.g1_affine_from_mont_limbs(sample[2]) == O.srs_point(t, n - 1)
    srs.close()
