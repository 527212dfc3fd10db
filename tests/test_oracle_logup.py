"""The oracle's C restatement of the LogUp prover vectors (logup/prover.rs), pinned by
the reference's own unit-test vectors for compute_multiplicities (logup/prover.rs:481-628:
unique table, duplicated table values, blinding rows, a value missing from the table)
and, for compute_logderivative, by what the argument's soundness fixes: the helper
columns satisfy h * prod(f + beta) = sum of the other factors' products, the aggregator
starts at zero and returns to zero on the last usable row (the LogUp balance), and a
wrong multiplicity breaks that balance (CPU only)."""
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C

R = O.R_MOD


def rows(vals):
    return C.ints_to_fr_rows([v % R for v in vals])


def ints(a):
    return C.fr_rows_to_ints(a)


ONES4 = [1, 1, 1, 1]


def test_reference_multiplicities_unique_table():
    # logup/prover.rs:486-523: table [1,2,3,4], inputs [1,2,3,3] and [2,2,3,4] -> [1,3,3,1]
    m = C.logup_multiplicities(rows(ONES4), [rows([1, 2, 3, 3]), rows([2, 2, 3, 4])], rows([1, 2, 3, 4]), 4, [])
    assert ints(m) == [1, 3, 3, 1]


def test_reference_multiplicities_value_missing():
    # logup/prover.rs:525-547: input value 5 is not in the table -> panic
    with pytest.raises(ValueError, match="input 0 row 2"):
        C.logup_multiplicities(rows(ONES4), [rows([1, 2, 5, 3])], rows([1, 2, 3, 4]), 4, [])


def test_reference_multiplicities_duplicated_table_values():
    # logup/prover.rs:549-576: table [1,2,2,3], input [1,2,2,3] -> each 2 gets 2/2 = 1
    m = C.logup_multiplicities(rows(ONES4), [rows([1, 2, 2, 3])], rows([1, 2, 2, 3]), 4, [])
    assert ints(m) == [1, 1, 1, 1]


def test_reference_multiplicities_blinding_rows():
    # logup/prover.rs:578-605: 2 usable rows, blinding values 42, 43 on the others
    m = C.logup_multiplicities(rows(ONES4), [rows([1, 2, 999, 888])], rows([1, 2, 0, 0]), 2, rows([42, 43]))
    assert ints(m) == [1, 1, 42, 43]


def test_fractional_multiplicity_and_selector():
    # a value looked up 3 times that appears twice in the table: 3/2 on both rows; rows
    # whose selector is zero are not counted
    sel = [1, 1, 0, 1, 1, 1]
    m = C.logup_multiplicities(rows(sel), [rows([7, 7, 9, 7, 8, 8])], rows([7, 8, 7, 5, 9, 9]), 6, [])
    half = pow(2, -1, R)
    assert ints(m) == [3 * half % R, 2, 3 * half % R, 0, 0, 0]


def _instance(n, bf, ncols, seed):
    rng = random.Random(seed)
    usable = n - bf - 1
    table = [rng.randrange(R) for _ in range(usable // 3)]
    table = [table[i % len(table)] for i in range(usable)] + [rng.randrange(R) for _ in range(n - usable)]
    sel = [rng.choice([0, 1, 1]) for _ in range(n)]   # boolean, as selector expressions are
    inputs = [[rng.choice(table[:usable]) if i < usable else rng.randrange(R) for i in range(n)]
              for _ in range(ncols)]
    return table, sel, inputs, usable


@pytest.mark.parametrize("ncols,chunks", [(1, [1]), (3, [2, 1]), (4, [1, 3])])
def test_logderivative_balance(ncols, chunks):
    n, bf = 64, 5
    table, sel, inputs, usable = _instance(n, bf, ncols, 7 + ncols)
    rng = random.Random(3)
    blind_m = [rng.randrange(R) for _ in range(n - usable)]
    m = C.logup_multiplicities(rows(sel), [rows(c) for c in inputs], rows(table), usable, rows(blind_m))
    beta = rng.randrange(R)
    blind_z = [rng.randrange(R) for _ in range(bf)]
    helpers, z = C.logup_logderivative(rows(sel), m, rows(table), [rows(c) for c in inputs], chunks,
                                       rows([beta])[0], rows(blind_z), bf)
    zi = ints(z)
    assert zi[0] == 0                                  # l_0 Z = 0
    assert zi[usable] == 0                             # the balance on the last usable row
    assert zi[n - bf:] == blind_z
    # helper columns: h prod_c (f_c + beta) = sum_c prod_(c' != c) (f_c' + beta)
    col = 0
    for k, size in enumerate(chunks):
        h = ints(helpers[k])
        for i in range(n):
            fs = [(inputs[col + c][i] + beta) % R for c in range(size)]
            prod = 1
            for f in fs:
                prod = prod * f % R
            rhs = 0
            for c in range(size):
                p = 1
                for c2 in range(size):
                    if c2 != c:
                        p = p * fs[c2] % R
                rhs = (rhs + p) % R
            assert h[i] * prod % R == rhs, (k, i)
        col += size
    # a wrong multiplicity breaks the balance
    bad = m.copy()
    bad[3] = rows([ints(m[3:4])[0] + 1])[0]
    _, zb = C.logup_logderivative(rows(sel), bad, rows(table), [rows(c) for c in inputs], chunks,
                                  rows([beta])[0], rows(blind_z), bf)
    assert ints(zb)[usable] != 0
