"""Device fr / fp arithmetic and G1 formulas against the oracle (GPU)."""
import random

import numpy as np
import pytest

from oracle import bls12_381 as O
from tests.helpers import empty_dev, from_dev, mzk, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("field", [0, 1])
def test_field_ops_bit_exact(ctx, field):
    mod, nl, tm, fm = ((O.R_MOD, 4, O.fr_to_mont, O.fr_from_mont) if field == 0
                       else (O.P_MOD, 6, O.fp_to_mont, O.fp_from_mont))
    rng = random.Random(field)
    n = 512
    xs = [rng.randrange(mod) for _ in range(n)]
    ys = [rng.randrange(mod) for _ in range(n)]
    xs[0], ys[1], xs[2], ys[2], xs[3], ys[3] = 0, 0, mod - 1, mod - 1, 1, mod - 1
    A = to_dev([O.to_limbs(tm(x), nl) for x in xs])
    B = to_dev([O.to_limbs(tm(y), nl) for y in ys])
    out = empty_dev(n, nl)
    ops = {0: lambda x, y: x * y, 1: lambda x, y: x + y, 2: lambda x, y: x - y, 3: lambda x, y: x * x,
           5: lambda x, y: x * y}
    for op, fn in ops.items():
        mzk.diag_field_op(ctx, field, op, A, B, out, n)
        got = from_dev(out)
        for i in range(n):
            raw = O.from_limbs(int(v) for v in got[i])
            assert raw < mod, "result not fully reduced"
            assert fm(raw) == fn(xs[i], ys[i]) % mod, (op, i)
    mzk.diag_field_op(ctx, field, 4, A, B, out, n)
    got = from_dev(out)
    assert [O.from_limbs(int(v) for v in got[i]) for i in range(n)] == xs


def test_field_mul_reference_kats(ctx):
    import json, os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
    for field, key, nl in [(0, "fr", 4), (1, "fp", 6)]:
        a = [int(h, 16) for h in kat[key]["mul_a_mont"]]
        b = [int(h, 16) for h in kat[key]["mul_b_mont"]]
        ab = [int(h, 16) for h in kat[key]["mul_ab_mont"]]
        A, B, out = to_dev([a]), to_dev([b]), empty_dev(1, nl)
        mzk.diag_field_op(ctx, field, 0, A, B, out, 1)
        assert [int(v) for v in from_dev(out)[0]] == ab


def test_g1_formulas_with_exceptional_cases(ctx):
    rng = random.Random(5)
    m = 64
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(m)]
    qts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(m)]
    qts[0] = pts[0]               # doubling reached through addition
    qts[1] = O.g1_neg(pts[1])     # P + (-P)
    qts[2] = None                 # identity operand
    pts[3] = None
    pts[4] = qts[4] = None
    A = to_dev([O.g1_affine_mont_limbs(p) for p in pts])
    B = to_dev([O.g1_affine_mont_limbs(q) for q in qts])
    out = empty_dev(m, 18)
    expect = {
        0: [O.g1_add(p, q) for p, q in zip(pts, qts)],
        1: [O.g1_double(p) for p in pts],
        2: [O.g1_add(O.g1_double(p), O.g1_double(q)) for p, q in zip(pts, qts)],
    }
    for op in range(3):
        mzk.diag_g1_op(ctx, op, A, B, out, m)
        got = from_dev(out)
        assert [O.g1_jacobian_from_mont_limbs(got[i]) for i in range(m)] == expect[op], op


def test_g1_two_g_kat(ctx):
    import json, os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))["g1"]
    G = O.g1_affine_mont_limbs(O.G1_GEN)
    A = to_dev([G])
    out = empty_dev(1, 18)
    mzk.diag_g1_op(ctx, 1, A, A, out, 1)
    two = O.g1_jacobian_from_mont_limbs(from_dev(out)[0])
    assert O.fp_to_mont(two[0]) == O.from_limbs(int(h, 16) for h in kat["two_g_x_mont"])
    assert O.fp_to_mont(two[1]) == O.from_limbs(int(h, 16) for h in kat["two_g_y_mont"])


def test_seeded_stream_matches_oracle(ctx):
    out = empty_dev(300, 4)
    ctx.fr_from_seed_dev(99, 3, 300, out, offset=17)
    ctx.synchronize()
    got = from_dev(out)
    for i in range(300):
        assert O.fr_from_mont(O.from_limbs(int(v) for v in got[i])) == O.fr_from_seed(99, 3, 17 + i)
