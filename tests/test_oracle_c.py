"""Pins the oracle's C restatement (the CPU baseline) against the Python oracle,
brute force and closed forms (CPU only)."""
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C


@pytest.fixture(scope="module", autouse=True)
def _lib():
    C.build()
    C.set_threads(8)


def aff_rows_to_points(rows):
    return [O.g1_affine_from_mont_limbs(r) for r in np.asarray(rows).reshape(-1, 12)]


def test_seed_stream_matches_python():
    a = C.fr_from_seed(1234, 7, 50, offset=3)
    assert C.fr_rows_to_ints(a) == [O.fr_from_seed(1234, 7, 3 + i) for i in range(50)]


def test_booth_index_matches_python():
    rng = random.Random(5)
    for _ in range(200):
        s = rng.randrange(O.R_MOD)
        rep = O.fr_to_bytes_le(s)
        buf = np.frombuffer(rep, dtype=np.uint8).copy()
        c = rng.randrange(1, 17)
        w = rng.randrange(O.FR_BITS // c + 1)
        assert C.lib().cpu_ref_booth_index(w, c, buf.ctypes.data) == O.get_booth_index(w, c, rep)


def test_window_size_rule():
    for n in [1, 3, 4, 31, 32, 100, 8103, 8104, 1 << 16, 1 << 20]:
        assert C.lib().cpu_ref_window_size(n) == O.msm_window_size(n)


def test_srs_bases_match_python():
    tau = O.fr_from_seed(99, 1, 0)
    rows = C.srs_bases(C.ints_to_fr_rows([tau])[0], 20)
    pts = aff_rows_to_points(rows)
    for i in [0, 1, 2, 7, 19]:
        assert pts[i] == O.srs_point(tau, i)


def test_msm_parallel_small_vs_naive():
    rng = random.Random(8)
    n = 40
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(n)]
    sc = [rng.randrange(O.R_MOD) for _ in range(n)]
    sc[3] = 0
    bases = np.array([O.g1_affine_mont_limbs(P) for P in pts], dtype=np.uint64)
    scal = C.ints_to_fr_rows(sc)
    ref = O.msm_naive(sc, pts)
    for fn in (C.msm_parallel, C.msm_best, C.msm_specific):
        got = O.g1_jacobian_from_mont_limbs(fn(scal, bases))
        assert got == ref


def test_msm_best_batch_affine_path_closed_form():
    # n >= 8104 so that c >= 10 and msm_best takes the batch-affine scheduler path
    n = 9000
    tau = O.fr_from_seed(5, 1, 0)
    bases = C.srs_bases(C.ints_to_fr_rows([tau])[0], n)
    scal = C.fr_from_seed(5, 2, n)
    expect = O.msm_srs_closed_form(C.fr_rows_to_ints(scal), tau)
    assert O.g1_jacobian_from_mont_limbs(C.msm_best(scal, bases)) == expect
    assert O.g1_jacobian_from_mont_limbs(C.msm_parallel(scal, bases)) == expect


def test_msm_best_exceptional_cases():
    # few distinct bases and their negatives force doubling and P + (-P) inside
    # batch_add (msm.rs:88-105) and the Jacobian fallback buckets
    rng = random.Random(21)
    n = 8200
    distinct = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(3)]
    choice = [rng.randrange(6) for _ in range(n)]
    pts = [distinct[c % 3] if c < 3 else O.g1_neg(distinct[c % 3]) for c in choice]
    pts[10] = None                      # identity base (dropped by msm_best)
    sc = [rng.randrange(1 << 20) for _ in range(n)]   # small scalars: few windows, dense buckets
    sc[11] = 0
    sc[12] = O.R_MOD - 1
    totals = [0, 0, 0]
    for s, c, P in zip(sc, choice, pts):
        if P is None:
            continue
        totals[c % 3] += s if c < 3 else -s
    expect = O.g1_sum([O.g1_mul(distinct[i], totals[i] % O.R_MOD) for i in range(3)])
    bases = np.array([O.g1_affine_mont_limbs(P) for P in pts], dtype=np.uint64)
    scal = C.ints_to_fr_rows(sc)
    for fn in (C.msm_best, C.msm_parallel, C.msm_specific):
        assert O.g1_jacobian_from_mont_limbs(fn(scal, bases)) == expect


@pytest.mark.parametrize("log_n", list(range(1, 11)))
def test_ntt_matches_python(log_n):
    vals = O.fr_vector_from_seed(3, log_n, 1 << log_n)
    a = C.ints_to_fr_rows(vals)
    out = C.ntt(a, log_n)
    assert C.fr_rows_to_ints(out) == O.dft(vals, O.root_of_unity(log_n))
    assert C.fr_rows_to_ints(C.intt(out, log_n)) == vals


def test_ntt_2_16_horner_samples():
    log_n = 16
    a = C.fr_from_seed(17, 1, 1 << log_n)
    vals = C.fr_rows_to_ints(a)
    out = C.fr_rows_to_ints(C.ntt(a, log_n))
    w = O.root_of_unity(log_n)
    for j in [0, 1, 12345, (1 << 16) - 1]:
        assert out[j] == O.poly_eval(vals, pow(w, j, O.R_MOD))


@pytest.mark.parametrize("k,ext_k", [(1, 3), (3, 5), (6, 8), (8, 10)])
def test_coset_extension_matches_python(k, ext_k):
    vals = O.fr_vector_from_seed(4, k, 1 << k)
    ext = C.coeff_to_extended(C.ints_to_fr_rows(vals), k, ext_k)
    assert C.fr_rows_to_ints(ext) == O.coeff_to_extended(vals, ext_k)
    back = C.extended_to_coeff(ext, ext_k)
    assert C.fr_rows_to_ints(back) == vals + [0] * ((1 << ext_k) - (1 << k))
    div = C.divide_by_vanishing_poly(ext, k, ext_k)
    assert C.fr_rows_to_ints(div) == O.divide_by_vanishing_poly(O.coeff_to_extended(vals, ext_k), k, ext_k)


# ---- the prover's scalar-vector restatements, against integer arithmetic
def test_vector_ops_match_integers():
    R = O.R_MOD
    vals_a = [0, 1, 2, R - 1, R - 2, 12345] + [O.fr_from_seed(3, 1, i) for i in range(40)]
    vals_b = [0, R - 1, 2, R - 1, 3, 54321] + [O.fr_from_seed(3, 2, i) for i in range(40)]
    a, b = C.ints_to_fr_rows(vals_a), C.ints_to_fr_rows(vals_b)
    assert C.fr_rows_to_ints(C.fr_vec_op(0, a, b)) == [(x + y) % R for x, y in zip(vals_a, vals_b)]
    assert C.fr_rows_to_ints(C.fr_vec_op(1, a, b)) == [(x - y) % R for x, y in zip(vals_a, vals_b)]
    assert C.fr_rows_to_ints(C.fr_vec_op(2, a, b)) == [x * y % R for x, y in zip(vals_a, vals_b)]
    s = O.fr_from_seed(9, 9, 0)
    sl = O.to_limbs(O.fr_to_mont(s), 4)
    sm = np.array(sl, dtype=np.uint64)
    assert C.fr_rows_to_ints(C.fr_vec_scalar(0, a, sm)) == [x * s % R for x in vals_a]
    assert C.fr_rows_to_ints(C.fr_vec_scalar(1, a, sm)) == [(x + s) % R for x in vals_a]
    acc = b.copy()
    C.fr_vec_scalar(2, a, sm, acc)
    assert C.fr_rows_to_ints(acc) == [(y + x * s) % R for x, y in zip(vals_a, vals_b)]


def test_batch_invert_prefix_powers_match_integers():
    R = O.R_MOD
    vals = [O.fr_from_seed(4, 1, i) for i in range(70)]
    vals[5] = vals[33] = 0
    a = C.ints_to_fr_rows(vals)
    assert C.fr_rows_to_ints(C.fr_batch_invert(a)) == [pow(v, -1, R) if v else 0 for v in vals]
    acc, want = 1, []
    for v in vals:
        acc = acc * v % R
        want.append(acc)
    assert C.fr_rows_to_ints(C.fr_prefix_product(a)) == want
    x = O.fr_from_seed(4, 2, 0)
    xm = np.array(O.to_limbs(O.fr_to_mont(x), 4), dtype=np.uint64)
    assert C.fr_rows_to_ints(C.fr_powers(xm, 50)) == [pow(x, i, R) for i in range(50)]


@pytest.mark.parametrize("n", [1, 2, 3, 10, 100])
def test_kate_division_is_synthetic_division(n):
    """q = (a(X) - a(b)) / (X - b) as exact polynomial long division."""
    R = O.R_MOD
    vals = [O.fr_from_seed(5, n, i) for i in range(n)]
    b = O.fr_from_seed(5, 0, 7)
    bm = np.array(O.to_limbs(O.fr_to_mont(b), 4), dtype=np.uint64)
    q = C.fr_rows_to_ints(C.fr_kate_div(C.ints_to_fr_rows(vals), bm))
    assert len(q) == n - 1
    # long division from the top: r_{n-1} = a_{n-1}; q_{i-1} = r_i; r_{i-1} = a_{i-1} + b q_{i-1}
    want = [0] * (n - 1)
    carry = 0
    for i in range(n - 1, 0, -1):
        carry = (vals[i] + b * carry) % R
        want[i - 1] = carry
    assert q == want
    # and q(X)(X - b) + a(b) = a(X) coefficient by coefficient
    ab = O.poly_eval(vals, b)
    prod = [0] * n
    for i, c in enumerate(q):
        prod[i + 1] = (prod[i + 1] + c) % R
        prod[i] = (prod[i] - b * c) % R
    prod[0] = (prod[0] + ab) % R
    assert prod == [v % R for v in vals]


def test_bases_from_scalars_match_python():
    s = [O.fr_from_seed(6, 1, i) for i in range(20)] + [0, 1, O.R_MOD - 1]
    rows = C.g1_mul_generator_many(C.ints_to_fr_rows(s))
    assert aff_rows_to_points(rows) == [O.g1_mul(O.G1_GEN, v) for v in s]


def test_msm_serial_adds_into_accumulator():
    rng = random.Random(12)
    n = 30
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(n)]
    sc = [rng.randrange(O.R_MOD) for _ in range(n)]
    bases = np.array([O.g1_affine_mont_limbs(P) for P in pts], dtype=np.uint64)
    acc_pt = O.g1_mul(O.G1_GEN, 987654321)
    acc = C.msm_parallel(C.ints_to_fr_rows([987654321]), np.array([O.g1_affine_mont_limbs(O.G1_GEN)],
                                                                    dtype=np.uint64))
    assert O.g1_jacobian_from_mont_limbs(acc) == acc_pt
    # msm.rs:364-367 doubles the incoming accumulator c times per window before adding the
    # window's buckets: acc_out = [2^(c W)] acc_in + MSM, W = 8 (highest non-zero byte) / c + 1
    sc[5] = (1 << 200) + 3                   # the top byte among the scalars decides W
    out = C.msm_serial(C.ints_to_fr_rows(sc), bases, acc)
    c = O.msm_window_size(n)
    top = max((s.bit_length() + 7) // 8 for s in sc)
    W = top * 8 // c + 1
    assert O.g1_jacobian_from_mont_limbs(out) == O.g1_add(O.g1_mul(acc_pt, pow(2, c * W, O.R_MOD)),
                                                          O.msm_naive(sc, pts))
    # from the identity it is the MSM itself; all-zero scalars leave acc untouched (:360-362)
    zero = np.zeros(18, dtype=np.uint64)
    assert O.g1_jacobian_from_mont_limbs(C.msm_serial(C.ints_to_fr_rows(sc), bases, zero)) == \
        O.msm_naive(sc, pts)
    assert O.g1_jacobian_from_mont_limbs(C.msm_serial(C.ints_to_fr_rows([0] * n), bases, acc)) == acc_pt


def _dft_textbook(a, w):
    n = len(a)
    return [sum(a[i] * pow(w, i * j, O.R_MOD) for i in range(n)) % O.R_MOD for j in range(n)]


@pytest.mark.parametrize("log_n", [1, 2, 3, 5])
def test_twiddle_table_ffts_with_any_primitive_root(log_n):
    # omega^3 is another primitive 2^log_n-th root (3 is odd): the table forms follow it
    n = 1 << log_n
    w = pow(O.root_of_unity(log_n), 3, O.R_MOD)
    tw = C.compute_twiddles(C.ints_to_fr_rows([w])[0], log_n)
    assert C.fr_rows_to_ints(tw) == [pow(w, i, O.R_MOD) for i in range(n // 2)]
    rng = random.Random(log_n)
    a = [rng.randrange(O.R_MOD) for _ in range(n)]
    assert C.fr_rows_to_ints(C.best_fft_with_twiddles(C.ints_to_fr_rows(a), tw, log_n)) == _dft_textbook(a, w)
    for n_real in sorted({0, 1, max(1, n // 2), n // 2 + 1, n} - {n + 1}):
        if n_real > n:
            continue
        b = a[:n_real] + [0] * (n - n_real)
        got = C.fft_coeff_to_extended(C.ints_to_fr_rows(b), tw, log_n, n_real)
        assert C.fr_rows_to_ints(got) == _dft_textbook(b, w), n_real


def test_suffix_sum_and_delta_commitment_identity():
    # params.rs:47-57: delta bases are suffix sums; committing a.to_delta() against them
    # gives the commitment of a against the plain bases (sum_i b_i sum_(j>=i) G_j =
    # sum_j a_j G_j), for the single and the double delta encodings
    rng = random.Random(21)
    n = 24
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(n)]
    pts[7] = None                                           # an identity base
    pts[11] = O.g1_neg(pts[10])                             # a suffix that cancels
    rows = np.array([O.g1_affine_mont_limbs(P) if P is not None else [0] * 12 for P in pts], dtype=np.uint64)
    suf = C.g1_suffix_sum(rows)
    for i in range(n):
        expect = O.g1_sum([P for P in pts[i:] if P is not None])
        assert O.g1_affine_from_mont_limbs(suf[i]) == expect, i
    a = [rng.randrange(O.R_MOD) for _ in range(n)]
    a[3:9] = [a[3]] * 6                                     # a constant run: zero deltas
    d1 = C.fr_rows_to_ints(C.fr_to_delta(C.ints_to_fr_rows(a), 1))
    assert d1 == [a[0]] + [(a[i] - a[i - 1]) % O.R_MOD for i in range(1, n)]
    assert d1[4:9] == [0] * 5
    d2 = C.fr_rows_to_ints(C.fr_to_delta(C.ints_to_fr_rows(a), 2))
    assert d2 == [d1[0]] + [(d1[i] - d1[i - 1]) % O.R_MOD for i in range(1, n)]
    plain = O.msm_naive(a, [P for P in pts])
    assert O.msm_naive(d1, [O.g1_affine_from_mont_limbs(r) for r in suf]) == plain
    suf2 = C.g1_suffix_sum(suf)
    assert O.msm_naive(d2, [O.g1_affine_from_mont_limbs(r) for r in suf2]) == plain
