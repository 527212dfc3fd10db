"""The oracle's G1 point decoders (ZCash serialisation, as ParamsKZG::read_custom reads
SRS files) pinned on the CPU: the published compressed generator, round trips through
the encoders for random points and their negatives, and each rejection class on a
point built for it (flags, x >= p, no square root, off the curve, outside G1)."""
import random

import pytest

from oracle import bls12_381 as O

# ZCash BLS12-381 serialisation: the compressed G1 generator
GEN_COMPRESSED = bytes.fromhex(
    "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac58"
    "6c55e83ff97a1aeffb3af00adb22c6bb")


def non_g1_point(start: int):
    """The first curve point with x >= start outside G1 (almost every curve point is:
    the cofactor exceeds 2^125)."""
    x = start
    while True:
        rhs = (x ** 3 + 4) % O.P_MOD
        y = pow(rhs, (O.P_MOD + 1) // 4, O.P_MOD)
        if y * y % O.P_MOD == rhs and not O.g1_in_subgroup((x, y)):
            return (x, y)
        x += 1


def test_generator_kat():
    assert O.g1_from_compressed(GEN_COMPRESSED) == O.G1_GEN
    assert O.g1_to_compressed(O.G1_GEN) == GEN_COMPRESSED
    assert O.g1_from_uncompressed(O.g1_to_uncompressed(O.G1_GEN)) == O.G1_GEN


def test_round_trips():
    rng = random.Random(5)
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(12)]
    pts += [O.g1_neg(P) for P in pts] + [None]
    for P in pts:
        c = O.g1_to_compressed(P)
        u = O.g1_to_uncompressed(P)
        assert O.g1_from_compressed(c) == P
        assert O.g1_from_uncompressed(u) == P
        assert O.g1_from_uncompressed(u, check_subgroup=False) == P
    # P and -P differ only in the sign flag
    P = pts[0]
    a, b = O.g1_to_compressed(P), O.g1_to_compressed(O.g1_neg(P))
    assert a[1:] == b[1:] and (a[0] ^ b[0]) == 0x20


def reason(fn, *args, **kw):
    with pytest.raises(O.PointDecodeError) as e:
        fn(*args, **kw)
    return e.value.reason


def test_rejections():
    P = O.g1_mul(O.G1_GEN, 12345)
    c = bytearray(O.g1_to_compressed(P))
    u = bytearray(O.g1_to_uncompressed(P))
    # flags
    assert reason(O.g1_from_compressed, bytes([c[0] & 0x7F]) + bytes(c[1:])) == "encoding"
    assert reason(O.g1_from_uncompressed, bytes([u[0] | 0x80]) + bytes(u[1:])) == "encoding"
    assert reason(O.g1_from_uncompressed, bytes([u[0] | 0x20]) + bytes(u[1:])) == "encoding"
    assert reason(O.g1_from_compressed, bytes([0xE0]) + bytes(47)) == "encoding"      # identity with sign
    assert reason(O.g1_from_compressed, bytes([0xC0]) + bytes(46) + b"\x01") == "encoding"
    # x >= p
    big = bytearray(O.P_MOD.to_bytes(48, "big"))
    big[0] |= 0x80
    assert reason(O.g1_from_compressed, bytes(big)) == "encoding"
    # no square root: the first x whose x^3 + 4 is a non-residue
    x = 1
    while pow((x ** 3 + 4) % O.P_MOD, (O.P_MOD - 1) // 2, O.P_MOD) == 1:
        x += 1
    nb = bytearray(x.to_bytes(48, "big"))
    nb[0] |= 0x80
    assert reason(O.g1_from_compressed, bytes(nb)) == "curve"
    # off the curve
    bad = O.g1_to_uncompressed((P[0], (P[1] + 1) % O.P_MOD))
    assert reason(O.g1_from_uncompressed, bad) == "curve"
    # outside G1: the order-3 point (0, 2) and a point of large order
    for Q in [(0, 2), non_g1_point(5)]:
        assert O.g1_is_on_curve(Q) and not O.g1_in_subgroup(Q)
        assert reason(O.g1_from_compressed, O.g1_to_compressed(Q)) == "subgroup"
        assert reason(O.g1_from_uncompressed, O.g1_to_uncompressed(Q)) == "subgroup"
        assert O.g1_from_uncompressed(O.g1_to_uncompressed(Q), check_subgroup=False) == Q
        assert O.g1_from_compressed(O.g1_to_compressed(Q), check_subgroup=False) == Q
