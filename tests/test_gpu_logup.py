"""LogUp prover vectors on the device (logup/prover.rs: compute_multiplicities and
compute_logderivative) against the oracle's C restatement, bit-exact: tables with
repeated values, several input columns in chunks, zero selector rows, blinding rows, a
value missing from the table reported with its column and row; and the running sums the
aggregator is built from (GPU)."""
import numpy as np
import pytest

from oracle import cpu_ref as C
from tests.helpers import empty_dev, from_dev, mzk, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def _instance(n, bf, ncols, seed):
    rng = np.random.default_rng(seed)
    usable = n - bf - 1
    distinct = C.fr_from_seed(seed, 1, max(1, usable // 3))
    table = np.concatenate([distinct[rng.integers(0, len(distinct), usable)],
                            C.fr_from_seed(seed, 2, n - usable)])
    one = C.fr_powers(C.fr_from_seed(seed, 3, 1)[0], 1)[0]
    sel = np.zeros((n, 4), dtype=np.uint64)
    sel[rng.random(n) < 0.7] = one
    inputs = [np.ascontiguousarray(np.concatenate([table[:usable][rng.integers(0, usable, usable)],
                                                   C.fr_from_seed(seed, 10 + c, n - usable)]))
              for c in range(ncols)]
    return np.ascontiguousarray(table), sel, inputs, usable


@pytest.mark.parametrize("log_n,ncols,chunks", [(6, 1, [1]), (10, 3, [2, 1]), (16, 4, [1, 3]), (18, 2, [2])])
def test_logup_matches_oracle(ctx, log_n, ncols, chunks):
    n, bf = 1 << log_n, 6
    print(f"\n[logup 2^{log_n}, {ncols} inputs]", flush=True)
    table, sel, inputs, usable = _instance(n, bf, ncols, log_n)
    bl_m = C.fr_from_seed(5, log_n, n - usable)
    m_ref = C.logup_multiplicities(sel, inputs, table, usable, bl_m)
    d_sel, d_tab = to_dev(sel), to_dev(table)
    d_in = [to_dev(x) for x in inputs]
    d_m = empty_dev(n, 4)
    ctx.logup_multiplicities(d_sel, d_in, d_tab, usable, bl_m, d_m)
    ctx.synchronize()
    assert np.array_equal(from_dev(d_m), m_ref)
    beta = C.fr_from_seed(6, log_n, 1)[0]
    bl_z = C.fr_from_seed(7, log_n, bf)
    h_ref, z_ref = C.logup_logderivative(sel, m_ref, table, inputs, chunks, beta, bl_z, bf)
    d_h = [empty_dev(n, 4) for _ in chunks]
    d_z = empty_dev(n, 4)
    ctx.logup_logderivative(d_sel, d_m, d_tab, d_in, chunks, beta, bl_z, d_h, d_z)
    ctx.synchronize()
    for dh, hr in zip(d_h, h_ref):
        assert np.array_equal(from_dev(dh), hr)
    z = from_dev(d_z)
    assert np.array_equal(z, z_ref)
    assert not z[0].any() and not z[usable].any()         # l_0 Z = 0 and the balance


def test_logup_missing_value_reported(ctx):
    n, bf = 256, 6
    table, sel, inputs, usable = _instance(n, bf, 2, 11)
    sel[40] = C.fr_powers(C.fr_from_seed(1, 1, 1)[0], 1)[0]
    inputs[1][40] = C.fr_from_seed(99, 99, 1)[0]             # not in the table, selected
    with pytest.raises(ValueError, match="input 1 row 40"):
        C.logup_multiplicities(sel, inputs, table, usable, C.fr_from_seed(5, 5, n - usable))
    d_m = empty_dev(n, 4)
    with pytest.raises(mzk.MzkError, match="input 1 row 40"):
        ctx.logup_multiplicities(to_dev(sel), [to_dev(x) for x in inputs], to_dev(table), usable,
                                 C.fr_from_seed(5, 5, n - usable), d_m)


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, (1 << 20) + 5])
def test_prefix_sums(ctx, n):
    a = C.fr_from_seed(12, n, n)
    d, out = to_dev(a), empty_dev(n, 4)
    ctx.fr_prefix_sum(d, out)
    ctx.synchronize()
    got = from_dev(out)
    # the running sums' differences give back the input (checked with the C vector ops)
    assert np.array_equal(got[:1], a[:1])
    if n > 1:
        assert np.array_equal(C.fr_vec_op(1, got[1:], got[:-1]), a[1:])
