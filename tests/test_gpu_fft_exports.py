"""Transforms above 2^22 (three-pass plans) and the fft.rs / msm.rs entry points
compute_twiddles, best_fft_with_twiddles, fft_coeff_to_extended, msm_parallel and
msm_serial, each against the oracle's C restatement, with Horner samples of the input
polynomial as an independent check at the large sizes (GPU)."""
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


def _root(log_n, power=1):
    return C.ints_to_fr_rows([pow(O.root_of_unity(log_n), power, O.R_MOD)])[0]


def _horner_samples(coeffs, outs, x0_int, w_int, idx):
    """outs[j] == sum_i coeffs[i] (x0 w^j)^i for the sampled j (C Horner evaluation)."""
    for j in idx:
        x = C.ints_to_fr_rows([x0_int * pow(w_int, j, O.R_MOD) % O.R_MOD])[0]
        assert np.array_equal(outs[j], C.poly_eval(coeffs, x)), j


@pytest.mark.parametrize("log_n", [23, 24])
def test_three_pass_forward_and_inverse(ctx, log_n):
    print(f"\n[ntt 2^{log_n} forward + inverse]", flush=True)
    n = 1 << log_n
    vals = C.fr_from_seed(23, log_n, n)
    d = to_dev(vals)
    out = empty_dev(n, 4)
    ctx.ntt(mzk.NttJob.make(d, out, log_n))
    ctx.synchronize()
    got = from_dev(out)
    ref = C.ntt(vals, log_n)
    assert np.array_equal(got, ref)
    _horner_samples(vals, got, 1, O.root_of_unity(log_n), [0, 1, 12345, n - 1])
    # inverse, in place, back to the coefficients
    ctx.ntt(mzk.NttJob.make(out, out, log_n, mzk.NTT_INVERSE))
    ctx.synchronize()
    assert np.array_equal(from_dev(out), vals)


def test_domain_extended_k_24(ctx):
    # EvaluationDomain::new(5, 22): extended_k = 24 (domain.rs:64-67)
    print("\n[domain (5, 22): coeff_to_extended 2^22 -> 2^24, extended_to_coeff]", flush=True)
    dom = ctx.domain(5, 22)
    assert (dom.k, dom.extended_k) == (22, 24)
    coeffs = C.fr_from_seed(24, 5, 1 << 22)
    din = to_dev(coeffs)
    ext = empty_dev(1 << 24, 4)
    dom.coeff_to_extended(din, ext)
    ctx.synchronize()
    got = from_dev(ext)
    assert np.array_equal(got, C.coeff_to_extended(coeffs, 22, 24))
    _horner_samples(coeffs, got, O.FR_ZETA, O.root_of_unity(24), [0, 7, (1 << 24) - 3])
    dom.extended_to_coeff(ext)
    ctx.synchronize()
    back = from_dev(ext)
    assert np.array_equal(back[:1 << 22], coeffs) and not back[1 << 22:].any()
    dom.close()


def test_three_pass_padded_coset_and_batch(ctx):
    # a padded coset transform (n_real not a power of two) of 2^23 in a batch with a
    # one-pass and a two-pass job, and a 2^23 inverse
    print("\n[ntt batch: 2^23 coset n_real=5e6, 2^12, 2^18 coset, 2^23 inverse]", flush=True)
    specs = [(23, mzk.NTT_COSET, 5_000_000), (12, 0, 4096), (18, mzk.NTT_COSET, 1 << 18),
             (23, mzk.NTT_INVERSE, 1 << 23)]
    outs, jobs, ins = [], [], []
    for i, (log_n, flags, n_real) in enumerate(specs):
        vals = C.fr_from_seed(400 + i, log_n, n_real)
        din = to_dev(vals)
        dout = empty_dev(1 << log_n, 4)
        ins.append((vals, din))
        outs.append(dout)
        jobs.append(mzk.NttJob.make(din, dout, log_n, flags, n_real))
    ctx.ntt_batch(jobs)
    ctx.synchronize()
    for (log_n, flags, n_real), (vals, _), dout in zip(specs, ins, outs):
        got = from_dev(dout)
        if flags & mzk.NTT_INVERSE:
            expect = C.intt(vals, log_n)
        else:
            padded = np.zeros((1 << log_n, 4), dtype=np.uint64)
            padded[:n_real] = vals
            expect = C.coeff_to_extended(padded, log_n, log_n) if flags & mzk.NTT_COSET else C.ntt(padded, log_n)
        assert np.array_equal(got, expect), (log_n, flags, n_real)
    vals0 = ins[0][0]
    _horner_samples(vals0, from_dev(outs[0]), O.FR_ZETA, O.root_of_unity(23), [3, (1 << 23) - 1])


def test_batch_rejects_overlapping_jobs(ctx):
    a = empty_dev(1 << 12, 4)
    b = empty_dev(1 << 12, 4)
    j_ok = mzk.NttJob.make(a, a, 12)                       # in place: allowed
    ctx.ntt_batch([j_ok, mzk.NttJob.make(b, b, 12)])
    j1 = mzk.NttJob.make(a, b, 12)
    j2 = mzk.NttJob.make(a[:2048], a[2048:], 11)           # its output is j1's input
    with pytest.raises(mzk.MzkError, match="overlaps"):
        ctx.ntt_batch([j1, j2])
    j4 = mzk.NttJob.make(a[2048:], b[2048:], 11)           # its output meets j1's output
    with pytest.raises(mzk.MzkError, match="overlaps"):
        ctx.ntt_batch([j1, j4])
    # inputs may be shared between jobs
    c1, c2 = empty_dev(1 << 12, 4), empty_dev(1 << 12, 4)
    ctx.ntt_batch([mzk.NttJob.make(a, c1, 12), mzk.NttJob.make(a, c2, 12), mzk.NttJob.make(b, b, 12)])
    ctx.synchronize()
    assert np.array_equal(from_dev(c1), from_dev(c2))


def test_ntt_sizes_beyond_the_field_rejected(ctx):
    a = empty_dev(16, 4)
    with pytest.raises(mzk.MzkError):
        ctx.ntt(mzk.NttJob.make(a, a, 33, 0, 16))
    with pytest.raises(mzk.MzkError):
        ctx.domain(5, 31)                                  # extended_k 33 > S = 32


@pytest.mark.parametrize("log_n", [1, 2, 5, 11, 16, 20, 23])
def test_twiddle_entry_points(ctx, log_n):
    # a primitive root other than the domain's (its cube), so nothing can come from the
    # library's own cached tables
    print(f"\n[fft exports 2^{log_n}]", flush=True)
    n = 1 << log_n
    w = _root(log_n, 3)
    tw = ctx.compute_twiddles(w, log_n)
    assert np.array_equal(tw, C.compute_twiddles(w, log_n))
    a = C.fr_from_seed(31, log_n, n)
    assert np.array_equal(ctx.best_fft_with_twiddles(a, tw, log_n), C.best_fft_with_twiddles(a, tw, log_n))
    assert np.array_equal(ctx.best_fft(a, w, log_n), C.best_fft_with_twiddles(a, tw, log_n))
    for n_real in sorted({1, max(1, n // 4), n // 2 + 1, n} & set(range(1, n + 1))):
        b = a.copy()
        b[n_real:] = 0
        got = ctx.fft_coeff_to_extended(b, tw, log_n, n_real)
        assert np.array_equal(got, C.fft_coeff_to_extended(b, tw, log_n, n_real)), n_real
    # the device form
    d = to_dev(a)
    dout = empty_dev(n, 4)
    ctx.best_fft_dev(d, dout, w, log_n, n // 2 if n > 1 else 1)
    ctx.synchronize()
    half = a.copy()
    half[max(1, n // 2):] = 0
    assert np.array_equal(from_dev(dout), C.best_fft_with_twiddles(half, tw, log_n))


def test_twiddle_table_checked(ctx):
    log_n = 6
    tw = C.compute_twiddles(_root(log_n), log_n)
    a = C.fr_from_seed(1, 1, 1 << log_n)
    bad = tw.copy()
    bad[0] = bad[1]                                        # twiddles[0] must be 1
    with pytest.raises(mzk.MzkError, match="twiddles"):
        ctx.best_fft_with_twiddles(a, bad, log_n)
    bad = C.compute_twiddles(_root(log_n + 1), log_n)      # a root of the wrong order
    with pytest.raises(mzk.MzkError, match="primitive"):
        ctx.best_fft_with_twiddles(a, bad, log_n)
    bad = tw.copy()
    bad[-1] = bad[-2]
    with pytest.raises(mzk.MzkError, match="twiddles"):
        ctx.fft_coeff_to_extended(a, bad, log_n, 10)


@pytest.mark.parametrize("n", [1, 3, 20, 100, 5000])
def test_msm_parallel_and_serial(ctx, n):
    tau = C.fr_from_seed(5, 0, 1)[0]
    bases = C.srs_bases(tau, n)
    sc = C.fr_from_seed(6, n, n)
    aff = O.g1_jacobian_from_mont_limbs
    assert aff(ctx.msm_parallel(sc, bases)) == aff(C.msm_parallel(sc, bases))
    # msm_serial doubles a non-identity accumulator c per window before adding (msm.rs:364)
    acc = C.msm_parallel(C.fr_from_seed(7, 7, 1), C.srs_bases(tau, 1))
    assert aff(ctx.msm_serial(sc, bases, acc)) == aff(C.msm_serial(sc, bases, acc))
    zero = np.zeros(18, dtype=np.uint64)
    assert aff(ctx.msm_serial(sc, bases, zero)) == aff(C.msm_parallel(sc, bases))
    zs = np.zeros_like(sc)
    assert aff(ctx.msm_serial(zs, bases, acc)) == aff(acc)
