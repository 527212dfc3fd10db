"""Lagrange and delta bases built on the device: g_to_lagrange (inverse DFT over G1
of the monomial bases) equals the bases [L_i(tau)] G of unsafe_setup computed from tau,
and suffix sums equal [sum_(j >= i) s_j] G, both checked point for point against the
oracle's fixed-base multiplication of scalars computed with Python integers (GPU)."""
import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import mzk, pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk
R = O.R_MOD


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def lagrange_scalars(tau, k):
    n = 1 << k
    w = O.root_of_unity(k)
    c = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    out = []
    for i in range(n):
        wi = pow(w, i, R)
        out.append(c * wi % R * pow((tau - wi) % R, -1, R) % R)
    return out


@pytest.mark.parametrize("k", [0, 1, 2, 3, 5, 8, 10])
def test_g_to_lagrange_matches_unsafe_setup(ctx, k):
    tau = 0x1234567890ABCDEF1234567 + k
    g = ctx.srs_generate(plonk.mont_limbs(tau), max(1, 1 << k) + 3)   # longer than 2^k: uses the prefix
    lag = g.to_lagrange(k)
    assert lag.n == 1 << k
    want = C.g1_mul_generator_many(C.ints_to_fr_rows(lagrange_scalars(tau, k)))
    assert np.array_equal(lag.download(), want)
    g.close()
    lag.close()


def test_g_to_lagrange_2p14_sampled(ctx):
    k = 14
    tau = 0xFEEDFACE ** 3 % R
    g = ctx.srs_generate(plonk.mont_limbs(tau), 1 << k)
    lag = g.to_lagrange(k)
    got = lag.download()
    idx = [0, 1, 2, 77, 4095, 8191, 12345, (1 << k) - 1]
    sc = lagrange_scalars_sample(tau, k, idx)
    want = C.g1_mul_generator_many(C.ints_to_fr_rows(sc))
    assert np.array_equal(got[idx], want)
    g.close()
    lag.close()


def lagrange_scalars_sample(tau, k, idx):
    n = 1 << k
    w = O.root_of_unity(k)
    c = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    return [c * pow(w, i, R) % R * pow((tau - pow(w, i, R)) % R, -1, R) % R for i in idx]


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 64 * 256 + 5, 70000])
def test_suffix_sum(ctx, n):
    s = [O.fr_from_seed(3, 1, i) for i in range(n)]
    if n > 5:
        s[3] = 0                                   # an identity base
    rows = C.ints_to_fr_rows(s)
    import torch
    base = ctx.srs_upload(C.g1_mul_generator_many(rows))
    suf = base.suffix_sum()
    acc, want = 0, [0] * n
    for i in range(n - 1, -1, -1):
        acc = (acc + s[i]) % R
        want[i] = acc
    assert np.array_equal(suf.download(), C.g1_mul_generator_many(C.ints_to_fr_rows(want)))
    base.close()
    suf.close()


def test_double_delta_commit_identity(ctx):
    """sum_j a_j L_j = sum_i d_i delta_i with d the first differences of a (the
    LagrangeDelta basis of proofs/src/poly/mod.rs): one commitment two ways."""
    k = 9
    n = 1 << k
    tau = 987654321
    g = ctx.srs_generate(plonk.mont_limbs(tau), n)
    lag = g.to_lagrange(k)
    delta = lag.suffix_sum()
    import torch
    a = [O.fr_from_seed(5, 2, i) for i in range(n)]
    d = [a[0]] + [(a[i] - a[i - 1]) % R for i in range(1, n)]
    ta = torch.from_numpy(C.ints_to_fr_rows(a).view(np.int64)).cuda()
    td = torch.from_numpy(C.ints_to_fr_rows(d).view(np.int64)).cuda()
    c1 = mzk.g1_to_affine(lag.commit(ta))
    c2 = mzk.g1_to_affine(delta.commit(td))
    assert np.array_equal(c1, c2)
    for x in (g, lag, delta):
        x.close()
