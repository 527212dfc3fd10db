"""The radix-2^28 field and XYZZ code of the MSM kernels, compiled for the host with
g++ and checked against the Python oracle: products and squares of random and edge
values, and mixed-addition chains (including doubling and P + (-P) inside a bucket),
XYZZ addition and doubling, compared as affine points (CPU only)."""
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from tests.helpers import BUILD_MODES, build_native

from oracle import bls12_381 as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", params=BUILD_MODES)
def exe(tmp_path_factory, request):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("f28_" + request.param) / "f28_check")
    return build_native(gxx, os.path.join(HERE, "native", "f28_check.cpp"), out, request.param)


def run(exe, cmd, n, payload: np.ndarray) -> np.ndarray:
    head = f"{cmd} {n}\n".encode()
    r = subprocess.run([exe], input=head + payload.astype("<u8").tobytes(), capture_output=True, check=True)
    return np.frombuffer(r.stdout, dtype="<u4")


def fp_limbs(v):            # canonical int -> blst Montgomery limbs
    return O.to_limbs(O.fp_to_mont(v), 6)


def fp_from_u32(words):     # 12 u32 Montgomery -> canonical int
    m = 0
    for i, w in enumerate(words):
        m |= int(w) << (32 * i)
    return O.fp_from_mont(m)


def test_products_and_squares(exe):
    rng = random.Random(3)
    P = O.P_MOD
    vals = [0, 1, P - 1, P - 2, (1 << 380), 2 ** 255 + 7] + [rng.randrange(P) for _ in range(300)]
    pairs = [(a, vals[(i * 7 + 3) % len(vals)]) for i, a in enumerate(vals)]
    pay = np.array([fp_limbs(a) + fp_limbs(b) for a, b in pairs], dtype=np.uint64)
    out = run(exe, "mul", len(pairs), pay).reshape(-1, 12)
    assert [fp_from_u32(r) for r in out] == [a * b % P for a, b in pairs]
    pay = np.array([fp_limbs(a) for a in vals], dtype=np.uint64)
    out = run(exe, "sqr", len(vals), pay).reshape(-1, 12)
    assert [fp_from_u32(r) for r in out] == [a * a % P for a in vals]


def xyzz_to_affine(words):
    X, Y, ZZ, ZZZ = (fp_from_u32(words[12 * i:12 * i + 12]) for i in range(4))
    if ZZ == 0:
        return None
    P = O.P_MOD
    return (X * pow(ZZ, -1, P) % P, Y * pow(ZZZ, -1, P) % P)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mixed_addition_chains(exe, seed):
    rng = random.Random(seed)
    base = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(8)]
    pts, signs, acc = [], [], None
    for i in range(120):
        if i % 17 == 5 and acc is not None:
            q, s = acc, rng.randrange(2)          # the running sum itself: doubling or P + (-P)
        elif i % 29 == 11:
            q, s = None, 0                        # the affine identity
        else:
            q, s = base[rng.randrange(8)], rng.randrange(2)
        pts.append(q)
        signs.append(s)
        if q is not None:
            acc = O.g1_add(acc, O.g1_neg(q) if s else q)
    pay = np.concatenate([np.array([O.g1_affine_mont_limbs(q) for q in pts], dtype=np.uint64).reshape(-1),
                          np.array(signs, dtype=np.uint64)])
    out = run(exe, "madd", len(pts), pay)
    assert xyzz_to_affine(out) == acc


def test_xyzz_add_and_double(exe):
    rng = random.Random(7)
    base = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(6)]
    for trial in range(4):
        pts = [base[rng.randrange(6)] for _ in range(20)]
        signs = [rng.randrange(2) for _ in range(20)]
        if trial == 1:                            # the two chains end equal: P + P
            pts[1::2] = pts[0::2]
            signs[1::2] = signs[0::2]
        if trial == 2:                            # the chains are negatives: P + (-P)
            pts[1::2] = pts[0::2]
            signs[1::2] = [1 - s for s in signs[0::2]]
        total = None
        for q, s in zip(pts, signs):
            total = O.g1_add(total, O.g1_neg(q) if s else q)
        pay = np.concatenate([np.array([O.g1_affine_mont_limbs(q) for q in pts], dtype=np.uint64).reshape(-1),
                              np.array(signs, dtype=np.uint64)])
        out = run(exe, "add", len(pts), pay)
        assert xyzz_to_affine(out[:48]) == total
        assert xyzz_to_affine(out[48:]) == O.g1_add(total, total)
