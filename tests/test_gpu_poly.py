"""Scalar-vector kernels of the prover (vector arithmetic, batch inversion, prefix
products, Horner evaluation, Kate division, powers, bases from scalars) on the device
against the oracle's C restatement, bit-exact, at sizes that cross every tile and
chunk boundary of the kernels (GPU)."""
import ctypes

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import from_dev, mzk, pkg, to_dev

pytestmark = pytest.mark.gpu
plonk = pkg.plonk


@pytest.fixture(scope="module")
def be():
    C.set_threads(16)
    b = plonk.GpuBackend()
    yield b
    b.ctx.close()


def host(be, v):
    be.sync()
    return from_dev(v).reshape(-1, 4)


SIZES = [1, 2, 31, 32, 33, 63, 64, 65, 4095, 4096, 4097, 65537, (1 << 20) + 3]


@pytest.mark.parametrize("n", SIZES)
def test_vec_ops(be, n):
    a = C.fr_from_seed(1, n, n)
    b = C.fr_from_seed(2, n, n)
    da, db = to_dev(a), to_dev(b)
    for op, f in ((0, be.add), (1, be.sub), (2, be.mul)):
        assert np.array_equal(host(be, f(da, db)), C.fr_vec_op(op, a, b)), op
    s = 0x1234567890ABCDEF1234567890ABCDEF % plonk.R_MOD
    sl = plonk.mont_limbs(s)
    assert np.array_equal(host(be, be.scale(da, s)), C.fr_vec_scalar(0, a, sl))
    assert np.array_equal(host(be, be.add_scalar(da, s)), C.fr_vec_scalar(1, a, sl))
    acc = C.fr_from_seed(3, n, n)
    want = C.fr_vec_scalar(2, a, sl, acc.copy())
    assert np.array_equal(host(be, be.axpy(to_dev(acc), da, s)), want)


def test_vec_ops_small_values_by_integers(be):
    a = [0, 1, 2, plonk.R_MOD - 1, 12345, plonk.R_MOD - 2]
    b = [0, plonk.R_MOD - 1, 2, plonk.R_MOD - 1, 54321, 3]
    da, db = to_dev(C.ints_to_fr_rows(a)), to_dev(C.ints_to_fr_rows(b))
    R = plonk.R_MOD
    assert C.fr_rows_to_ints(host(be, be.add(da, db))) == [(x + y) % R for x, y in zip(a, b)]
    assert C.fr_rows_to_ints(host(be, be.sub(da, db))) == [(x - y) % R for x, y in zip(a, b)]
    assert C.fr_rows_to_ints(host(be, be.mul(da, db))) == [x * y % R for x, y in zip(a, b)]


# tiles of 1024 values (256 threads x 4): one partial tile, tile edges, and more tiles than
# the totals' workgroup has threads (256 x 1024 = 262144 values)
@pytest.mark.parametrize("n", [1, 5, 32, 33, 1000, 1023, 1024, 1025, 65536 + 7, 262144 + 1024 * 3 + 1,
                               (1 << 20) + 1])
def test_batch_invert(be, n):
    a = C.fr_from_seed(4, n, n)
    a[::7] = 0                       # zeros stay zero (BatchInvert skips them)
    got = host(be, be.batch_invert(to_dev(a)))
    assert np.array_equal(got, C.fr_batch_invert(a))
    if n <= 1000:
        ints = C.fr_rows_to_ints(a)
        assert C.fr_rows_to_ints(got) == [pow(v, -1, O.R_MOD) if v else 0 for v in ints]


def test_batch_invert_all_zero_and_single(be):
    z = np.zeros((3000, 4), dtype=np.uint64)
    assert not host(be, be.batch_invert(to_dev(z))).any()
    one = C.fr_from_seed(9, 1, 1)
    assert C.fr_rows_to_ints(host(be, be.batch_invert(to_dev(one)))) == [pow(C.fr_rows_to_ints(one)[0], -1, O.R_MOD)]


@pytest.mark.parametrize("n", [1, 2, 255, 256, 4095, 4096, 4097, 8192 * 3 + 5, (1 << 18) + 1, (1 << 20)])
def test_prefix_product(be, n):
    a = C.fr_from_seed(5, n, n)
    got = host(be, be.prefix_product(to_dev(a)))
    assert np.array_equal(got, C.fr_prefix_product(a))


def test_prefix_product_batch_equals_per_call(be):
    L = mzk.load_library()
    sizes = [1, 2, 4095, 4096, 4097, 0, 3 * 4096 + 5, 300, (1 << 17) + 3]
    ins = [to_dev(C.fr_from_seed(60 + i, n, n)) if n else None for i, n in enumerate(sizes)]
    outs = [be._empty(n) for n in sizes]
    jobs = (mzk.PrefixJob * len(sizes))()
    for j, (a, o, n) in enumerate(zip(ins, outs, sizes)):
        jobs[j].d_a = a.data_ptr() if n else None
        jobs[j].d_out = o.data_ptr() if n else None
        jobs[j].n = n
    print(f"\n[prefix batch] {len(sizes)} jobs", flush=True)
    assert L.mzk_fr_prefix_product_batch(be.ctx.h, ctypes.addressof(jobs), len(sizes)) == 0
    for a, o, n in zip(ins, outs, sizes):
        if n:
            assert np.array_equal(host(be, o), C.fr_prefix_product(host(be, a)))
    # an output inside another job's input is refused before anything runs
    jobs[1].d_out = ins[2].data_ptr()
    assert L.mzk_fr_prefix_product_batch(be.ctx.h, ctypes.addressof(jobs), len(sizes)) == 1


def test_prefix_product_integers(be):
    vals = [2, 3, 5, 7, 11, 0, 13]
    got = C.fr_rows_to_ints(host(be, be.prefix_product(to_dev(C.ints_to_fr_rows(vals)))))
    want, acc = [], 1
    for v in vals:
        acc = acc * v % O.R_MOD
        want.append(acc)
    assert got == want


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 16384, 16385, (1 << 18) + 9, 1 << 20])
def test_eval(be, n):
    a = C.fr_from_seed(6, n, n)
    x = 0xDEADBEEF0123456789 % O.R_MOD
    got = be.eval_many([(to_dev(a), x)])[0]
    want = plonk.from_mont_limbs(C.poly_eval(a, plonk.mont_limbs(x)))
    assert got == want
    if n <= 65:
        assert got == O.poly_eval(C.fr_rows_to_ints(a), x)


def test_eval_empty_is_zero(be):
    out = np.ones(4, dtype=np.uint64)
    L = mzk.load_library()
    rc = L.mzk_fr_eval(be.ctx.h, None, 0, plonk.mont_limbs(3).ctypes.data, out.ctypes.data)
    assert rc == 0 and not out.any()


@pytest.mark.parametrize("n", [1, 2, 3, 64, 65, 66, 129, 64 * 256 + 1, 64 * 256 * 3 + 17, (1 << 20) + 1])
def test_kate_division(be, n):
    a = C.fr_from_seed(7, n, n)
    b = 0xABCDEF0123456789ABCDEF % O.R_MOD
    got = host(be, be.kate_div(to_dev(a), b)) if n > 1 else np.zeros((0, 4), dtype=np.uint64)
    want = C.fr_kate_div(a, plonk.mont_limbs(b))
    assert got.shape == want.shape and np.array_equal(got, want)
    if n > 1:
        # q(X) (X - b) + a(b) = a(X), checked at a random point
        z = 0x5A5A5A5A12345 % O.R_MOD
        qz = be.eval_many([(to_dev(got), z)])[0]
        ab = plonk.from_mont_limbs(C.poly_eval(a, plonk.mont_limbs(b)))
        az = plonk.from_mont_limbs(C.poly_eval(a, plonk.mont_limbs(z)))
        assert (qz * (z - b) + ab) % O.R_MOD == az


def test_kate_division_batch_equals_per_call(be):
    # more jobs than one launch takes, mixed sizes (n = 1 writes nothing), distinct points
    sizes = [2, 3, 65, 1, 4096, 4097, 64 * 256 * 3 + 17, 129, (1 << 16) + 1, 7, 300, 64 * 256 + 1]
    polys = [C.fr_from_seed(40 + i, n, n) for i, n in enumerate(sizes)]
    pts = [(0x1234567 * (i + 3) + 11) % O.R_MOD for i in range(len(sizes))]
    print(f"\n[kate batch] {len(sizes)} jobs", flush=True)
    got = be.kate_div_many([(to_dev(a), b) for a, b in zip(polys, pts)])
    for a, b, q in zip(polys, pts, got):
        want = C.fr_kate_div(a, plonk.mont_limbs(b))
        assert np.array_equal(host(be, q) if q.shape[0] else np.zeros((0, 4), dtype=np.uint64), want)


def test_kate_division_batch_rejects_overlap(be):
    L = mzk.load_library()
    d = to_dev(C.fr_from_seed(2, 1, 128))          # every span below stays inside it
    jobs = (mzk.KateJob * 2)()
    for j in jobs:
        j.d_a = d.data_ptr()
        j.n = 32
        j.b[:] = [int(w) for w in plonk.mont_limbs(5)]
    jobs[0].d_q = d.data_ptr() + 32 * 32          # disjoint from both inputs
    jobs[1].d_q = d.data_ptr() + 32 * 16          # inside the inputs
    assert L.mzk_fr_kate_div_batch(be.ctx.h, ctypes.addressof(jobs), 2) == 1
    jobs[1].d_q = d.data_ptr() + 32 * 40          # overlaps job 0's output
    assert L.mzk_fr_kate_div_batch(be.ctx.h, ctypes.addressof(jobs), 2) == 1
    jobs[1].n = 0
    assert L.mzk_fr_kate_div_batch(be.ctx.h, ctypes.addressof(jobs), 2) == 1


@pytest.mark.parametrize("n", [1, 7, 4096, (1 << 20) + 5])
def test_powers(be, n):
    x = plonk.omega_for(20)
    got = host(be, be.powers(x, n))
    assert np.array_equal(got, C.fr_powers(plonk.mont_limbs(x), n))


def test_bases_from_scalars(be):
    n = 3000
    s = C.fr_from_seed(8, 1, n)
    srs = be.srs_from_scalars(to_dev(s))
    got = srs.download()
    assert np.array_equal(got, C.g1_mul_generator_many(s))
    srs.close()


def test_bad_arguments_are_errors(be):
    L = mzk.load_library()
    assert L.mzk_fr_vec_op(None, 0, None, None, None, 4) == 1
    assert L.mzk_fr_vec_op(be.ctx.h, 7, None, None, None, 4) == 1
    assert L.mzk_fr_kate_div(be.ctx.h, None, 0, plonk.mont_limbs(1).ctypes.data, None) == 1
    d = to_dev(C.fr_from_seed(1, 1, 8))
    assert L.mzk_fr_kate_div(be.ctx.h, d.data_ptr(), 8, plonk.mont_limbs(1).ctypes.data, d.data_ptr()) == 1
    assert L.mzk_fr_prefix_product(be.ctx.h, d.data_ptr(), d.data_ptr(), (1 << 24) + 1) == 1


def test_eval_batch_mixed_sizes_and_points(be):
    sizes = [0, 1, 64, 65, 1000, 16384, 16385, (1 << 18) + 3, 7]
    polys = [C.fr_from_seed(9, i, n) if n else np.zeros((0, 4), dtype=np.uint64) for i, n in enumerate(sizes)]
    xs = [(0x1111 * (i + 3)) ** 5 % O.R_MOD for i in range(len(sizes))]
    got = be.eval_many([(to_dev(p), x) for p, x in zip(polys, xs)])
    want = [plonk.from_mont_limbs(C.poly_eval(p, plonk.mont_limbs(x))) if p.shape[0] else 0
            for p, x in zip(polys, xs)]
    assert got == want


@pytest.mark.parametrize("n,ncols", [(1, 1), (100, 3), (4097, 2), ((1 << 18) + 3, 4)])
def test_perm_products(be, n, ncols):
    """Both products of permutation/prover.rs:129-160 for two sets in one call, against
    the same products composed from the C restatement's vector operations."""
    import ctypes
    L = mzk.load_library()
    R = plonk.R_MOD
    w = C.fr_powers(plonk.mont_limbs(pow(7, (R - 1) >> 20, R)), n)
    beta, gamma = 0xBE7A % R, 0x6A33A % R
    sets, want, outs, keep = [], [], [], []
    for s in range(2):
        vals = [C.fr_from_seed(40 + 10 * s + j, n, n) for j in range(ncols)]
        sig = [C.fr_from_seed(41 + 10 * s + j, n, n) for j in range(ncols)]
        dbs = [beta * pow(O.FR_DELTA, 3 * s + j, R) % R for j in range(ncols)]
        ones = np.tile(plonk.mont_limbs(1), (n, 1))
        den, num = ones.copy(), ones.copy()
        g = np.tile(plonk.mont_limbs(gamma), (n, 1))
        for j in range(ncols):
            bs = C.fr_vec_scalar(0, sig[j], plonk.mont_limbs(beta))
            den = C.fr_vec_op(2, den, C.fr_vec_op(0, C.fr_vec_op(0, vals[j], g), bs))
            dw = C.fr_vec_scalar(0, w, plonk.mont_limbs(dbs[j]))
            num = C.fr_vec_op(2, num, C.fr_vec_op(0, C.fr_vec_op(0, vals[j], g), dw))
        want.append((num, den))
        e = mzk.PermSet()
        dv = [to_dev(v) for v in vals] + [to_dev(v) for v in sig]
        keep += dv
        for j in range(ncols):
            e.values[j] = dv[j].data_ptr()
            e.sigma[j] = dv[ncols + j].data_ptr()
            e.delta_beta[j][:] = [int(x) for x in plonk.mont_limbs(dbs[j])]
        e.beta[:] = [int(x) for x in plonk.mont_limbs(beta)]
        e.gamma[:] = [int(x) for x in plonk.mont_limbs(gamma)]
        e.ncols = ncols
        o_num, o_den = to_dev(np.zeros((n, 4), dtype=np.uint64)), to_dev(np.zeros((n, 4), dtype=np.uint64))
        outs.append((o_num, o_den))
        e.num, e.den = o_num.data_ptr(), o_den.data_ptr()
        sets.append(e)
    arr = (mzk.PermSet * 2)(*sets)
    dw_ = to_dev(w)
    be.sync()
    assert L.mzk_plonk_perm_products_dev(be.ctx.h, dw_.data_ptr(), n, ctypes.cast(arr, ctypes.c_void_p), 2) == 0
    for (o_num, o_den), (num, den) in zip(outs, want):
        assert np.array_equal(host(be, o_num), num) and np.array_equal(host(be, o_den), den)


@pytest.mark.parametrize("n", [1, 2, 3, 257, (1 << 18) + 1])
def test_delta_encodings(be, n):
    a = C.fr_from_seed(8, n, n)
    if n > 100:
        a[50:90] = a[50]                                  # a constant run
    d = to_dev(a)
    for order in (1, 2):
        assert np.array_equal(host(be, be.to_delta(d, order)), C.fr_to_delta(a, order)), order
    with pytest.raises(mzk.MzkError, match="overlaps"):
        be.mzk._check(be._L.mzk_fr_to_delta_dev(be.ctx.h, be._p(d), be._p(d), n, 1))


def test_delta_basis_commit_equals_lagrange_commit(be):
    # a z-like vector (constant runs with a few steps) committed as its deltas against the
    # suffix-sum bases gives its Lagrange commitment (permutation/prover.rs:236-245)
    k = 16
    params = plonk.Params(be, k, 0xDE17A)
    n = 1 << k
    rng = np.random.default_rng(3)
    steps = np.sort(rng.choice(n, 900, replace=False))
    vals = C.fr_from_seed(9, 9, len(steps) + 1)
    z = np.repeat(vals, np.diff(np.concatenate([[0], steps, [n]])), axis=0)
    dz = to_dev(np.ascontiguousarray(z))
    plain = be.commit_jobs([(params.g_lagrange, dz)])[0]
    delta = be.commit_jobs([(params.g_lagrange_delta, be.to_delta(dz))])[0]
    assert np.array_equal(plain, delta)
    double = params.g_lagrange_delta.suffix_sum()
    assert np.array_equal(be.commit_jobs([(double, be.to_delta(dz, 2))])[0], plain)
    # the delta bases themselves against the C restatement
    assert np.array_equal(params.g_lagrange_delta.download(), C.g1_suffix_sum(params.g_lagrange.download()))
