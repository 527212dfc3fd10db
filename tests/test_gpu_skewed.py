"""MSMs of skewed scalars (GPU): boolean, small-valued and constant columns put most window
digits into a few buckets, whose chunk partials the reductions then merge from one slot
after the long-bucket merge.  Every result against the C restatement of msm_best; each
scalar kind through the single shared set, a batch of three shared sets (block tree), a
batch of six (segments) and the per-window sets of host bases."""
import math
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import mzk, to_dev

pytestmark = pytest.mark.gpu

N_SRS = 1 << 16


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def srs(ctx):
    tau = C.fr_from_seed(77, 0, 1)[0]
    bases = C.srs_bases(tau, N_SRS)
    s = ctx.srs_upload(bases)
    yield bases, s
    s.close()


def aff(j):
    return O.g1_jacobian_from_mont_limbs(j)


def scalars(kind, n, seed):
    rng = random.Random(seed)
    if kind == "bool":
        vals = [rng.randrange(2) for _ in range(n)]
    elif kind == "nibble":
        vals = [rng.randrange(16) for _ in range(n)]
    elif kind == "const":
        vals = [5] * n
    else:   # one huge run of a repeated value among random ones
        vals = [rng.randrange(O.R_MOD) if i % 8 == 0 else 3 for i in range(n)]
    return C.ints_to_fr_rows(vals)


KINDS = ["bool", "nibble", "const", "runs"]


@pytest.mark.parametrize("kind", KINDS)
def test_single_shared_set(ctx, srs, kind):
    bases, s = srs
    n = 1 << 16
    sc = scalars(kind, n, 1)
    assert aff(s.commit(to_dev(sc), n)) == aff(C.msm_best(sc, bases[:n]))
    assert ctx.msm_last_plan()["bucket_sets"] == 1


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("nj", [3, 6])
def test_batch_shared_sets(ctx, srs, kind, nj):
    bases, s = srs
    n = (1 << 15) + 3
    rng = random.Random(nj)
    host, jobs = [], []
    for i in range(nj):
        sc = scalars(kind, n, 10 * nj + i)
        off = rng.randrange(0, N_SRS - n + 1)
        host.append((sc, off))
        jobs.append((to_dev(sc), n, off))
    outs = s.commit_batch(jobs)
    assert ctx.msm_last_plan()["bucket_sets"] == nj
    for (sc, off), out in zip(host, outs):
        assert aff(out) == aff(C.msm_best(sc, bases[off:off + n])), (kind, nj, off)


@pytest.mark.parametrize("kind", KINDS)
def test_per_window_sets(ctx, srs, kind):
    bases, _ = srs
    n = 1 << 14
    sc = scalars(kind, n, 3)
    out = ctx.msm_best(sc, bases[:n])
    assert ctx.msm_last_plan()["bucket_sets"] > 1
    assert aff(out) == aff(C.msm_best(sc, bases[:n]))


def test_long_bucket_threshold_edges(ctx, srs):
    """Buckets spanning just under, at and just over the long-bucket threshold (64 chunks):
    a column of ones and zeros, all of whose digits fall in one bucket.  The chunk size
    comes from the digits found (ceil(M / chunks), at least 2 and at least sqrt(D) / 4 for
    the largest bucket's D digits), so k c ones at the k the rule gives span c chunks."""
    bases, s = srs
    n = 1 << 16
    for chunks in (63, 64, 65, 66):
        # the chunk size k the rule gives k chunks' worth of ones: 4, 4, 5, 5
        k = next(k for k in range(2, 16) if max(2, math.ceil(math.sqrt(k * chunks) / 4)) == k)
        ones = k * chunks
        vals = [1] * ones + [0] * (n - ones)
        sc = C.ints_to_fr_rows(vals)
        assert aff(s.commit(to_dev(sc), n)) == aff(C.msm_best(sc, bases[:n])), chunks
        plan = ctx.msm_last_plan()
        assert plan["refs_per_thread"] == k and plan["digits"] == ones, plan


@pytest.mark.parametrize("density", [0.001, 0.05, 0.3])
def test_sparse_columns(ctx, srs, density):
    """Mostly-zero columns: the chunk size follows the digits found, not the column's
    length; single commit and a batch of three against msm_best."""
    bases, s = srs
    n = 1 << 16
    rng = random.Random(int(density * 1000))
    cols = []
    for i in range(3):
        vals = [rng.randrange(1, O.R_MOD) if rng.random() < density else 0 for _ in range(n)]
        cols.append(C.ints_to_fr_rows(vals))
    assert aff(s.commit(to_dev(cols[0]), n)) == aff(C.msm_best(cols[0], bases[:n]))
    outs = s.commit_batch([(to_dev(c), n, 0) for c in cols])
    for c, out in zip(cols, outs):
        assert aff(out) == aff(C.msm_best(c, bases[:n]))


@pytest.mark.parametrize("ones,pattern", [(120, "same"), (120, "alternate"), (1 << 16, "same")])
def test_repeated_points_in_one_bucket(ctx, ones, pattern):
    """A column of ones over repeated bases: 120 ones take 40 chunks of 3, whose partials
    are equal points (the quad addition's doubling) or, over alternating P and -P, opposite
    points (its cancellation to the identity); every other bucket is empty, so the channel
    trees add identity operands throughout.  2^16 ones of one point: the long-bucket merge
    of equal partials.  Against msm_best on the same bases."""
    tau = C.fr_from_seed(78, 0, 1)[0]
    n = 1 << 16
    bs = C.srs_bases(tau, n)
    p = bs[5].copy()
    neg = np.array(O.g1_affine_mont_limbs(O.g1_neg(O.g1_affine_from_mont_limbs(p))), dtype=np.uint64)
    for i in range(ones):
        bs[i] = neg if pattern == "alternate" and i % 2 else p
    s = ctx.srs_upload(bs)
    sc = C.ints_to_fr_rows([1] * ones + [0] * (n - ones))
    got = aff(s.commit(to_dev(sc), n))
    plan = ctx.msm_last_plan()
    s.close()
    assert got == aff(C.msm_best(sc, bs)), (ones, pattern, plan)
    if ones == 120:
        assert plan["refs_per_thread"] == 3 and plan["digits"] == ones, plan
