"""Pins the Python oracle against the reference's own known-answer vectors and
against what the mathematics fixes (CPU only)."""
import json
import os
import random

import pytest

from oracle import bls12_381 as O

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))


def L(hexlist):
    return O.from_limbs([int(h, 16) for h in hexlist])


# ---------------------------------------------------------------- fr (fq.rs)
def test_fr_modulus_and_inv():
    assert L(KAT["fr"]["modulus"]) == O.R_MOD
    # fq.rs:934-947 test_inv: -(r^-1) mod 2^64 by exponentiation
    inv = 1
    for _ in range(63):
        inv = inv * inv % (1 << 64)
        inv = inv * (O.R_MOD & O.M64) % (1 << 64)
    inv = (-inv) % (1 << 64)
    assert inv == int(KAT["fr"]["inv64"], 16)
    assert (O.R_MOD * inv + 1) % (1 << 64) == 0


def test_fr_montgomery_constants():
    f = KAT["fr"]
    assert O.fr_from_mont(L(f["R_mont"])) == 1
    assert L(f["R_mont"]) == O.FR_R
    assert L(f["R2_mont"]) == pow(2, 512, O.R_MOD)
    assert L(f["R3_mont"]) == pow(2, 768, O.R_MOD)
    assert O.fr_from_mont(L(f["generator_mont"])) == 7
    assert O.fr_from_mont(L(f["root_of_unity_mont"])) == O.FR_ROOT_OF_UNITY
    root = O.FR_ROOT_OF_UNITY
    assert pow(root, 1 << 32, O.R_MOD) == 1 and pow(root, 1 << 31, O.R_MOD) != 1
    assert O.fr_from_mont(L(f["root_of_unity_inv_mont"])) * root % O.R_MOD == 1
    assert O.fr_from_mont(L(f["two_inv_mont"])) == O.FR_TWO_INV
    assert O.fr_from_mont(L(f["delta_mont"])) == O.FR_DELTA
    assert O.fr_from_mont(L(f["zeta_mont"])) == O.FR_ZETA
    assert pow(O.FR_ZETA, 3, O.R_MOD) == 1 and O.FR_ZETA != 1


def test_fr_bytes_kats():
    f = KAT["fr"]
    # R2 as an element is 2^512 * R^-1 = 2^256 mod r
    assert list(O.fr_to_bytes_le(O.fr_from_mont(L(f["R2_mont"])))) == f["bytes_R2"]
    assert list(O.fr_to_bytes_le(O.R_MOD - 1)) == f["bytes_minus_one"]
    assert O.fr_from_bytes_le(bytes(f["bytes_minus_one"])) == O.R_MOD - 1
    assert O.fr_from_bytes_le(O.R_MOD.to_bytes(32, "little")) is None
    for bad in f["bytes_invalid"]:
        assert O.fr_from_bytes_le(bytes(bad)) is None
    # fq.rs:949-963 Debug is the big-endian hex of the canonical value
    assert f["debug_R2"] == "Fq(0x%064x)" % O.fr_from_mont(L(f["R2_mont"]))


def test_fr_arith_kats():
    f = KAT["fr"]
    largest = L(f["largest_raw"])
    assert (largest + largest) % O.R_MOD == L(f["largest_plus_largest_raw"])
    assert (largest + 1) % O.R_MOD == 0
    assert O.fr_mont_mul(L(f["mul_a_mont"]), L(f["mul_b_mont"])) == L(f["mul_ab_mont"])


def test_fr_from_uniform_bytes_matches_montgomery_recipe():
    # fq.rs:712-736 computes a0*R2 + a1*R3 with Montgomery products
    rng = random.Random(1)
    for _ in range(20):
        b = bytes(rng.getrandbits(8) for _ in range(64))
        a0 = int.from_bytes(b[:32], "little")
        a1 = int.from_bytes(b[32:], "little")
        mont = (O.fr_mont_mul(a0, pow(2, 512, O.R_MOD)) + O.fr_mont_mul(a1, pow(2, 768, O.R_MOD))) % O.R_MOD
        assert O.fr_from_mont(mont) == O.fr_from_uniform_bytes(b)


def test_splitmix64_known_answer():
    # first output of Vigna's splitmix64.c with state 0
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF


# ---------------------------------------------------------------- fp (fp.rs)
def test_fp_kats():
    f = KAT["fp"]
    assert L(f["modulus"]) == O.P_MOD
    assert L(f["R_mont"]) == O.FP_R
    assert L(f["R2_mont"]) == pow(2, 768, O.P_MOD)
    assert O.fp_mont_mul(L(f["mul_a_mont"]), L(f["mul_b_mont"])) == L(f["mul_ab_mont"])


# ---------------------------------------------------------------- G1 (g1.rs)
def test_g1_generator():
    g = KAT["g1"]
    G = (int(g["generator_x"]), int(g["generator_y"]))
    assert G == O.G1_GEN
    assert O.g1_is_on_curve(G)
    assert O.g1_mul(G, O.R_MOD) is None
    assert O.g1_mul(G, O.R_MOD - 1) == O.g1_neg(G)
    assert O.g1_to_compressed(G).hex() == g["compressed_generator"]


def test_g1_doubling_kat():
    g = KAT["g1"]
    two = O.g1_double(O.G1_GEN)
    assert O.fp_to_mont(two[0]) == L(g["two_g_x_mont"])
    assert O.fp_to_mont(two[1]) == L(g["two_g_y_mont"])
    assert O.g1_mul(O.G1_GEN, 2) == two


def test_g1_jacobian_scaled_generator_kat():
    # g1.rs:1042-1058: (x z^2, y z^3, z) is on the curve and equals G
    z = O.fp_from_mont(L(KAT["g1"]["z_mont"]))
    X = O.G1_GEN[0] * z * z % O.P_MOD
    Y = O.G1_GEN[1] * z * z * z % O.P_MOD
    assert O.jac_to_affine((X, Y, z)) == O.G1_GEN
    assert O.jac_to_affine(O._jac_add((X, Y, z), (1, 1, 0))) == O.G1_GEN


def test_g1_degenerate_addition_kat():
    g = KAT["g1"]
    beta = O.fp_from_mont(L(g["beta_mont"]))
    beta = beta * beta % O.P_MOD
    a = O.g1_mul(O.G1_GEN, 4)
    b = (a[0] * beta % O.P_MOD, (-a[1]) % O.P_MOD)
    assert O.g1_is_on_curve(b)
    c = O.g1_add(a, b)
    assert O.fp_to_mont(c[0]) == L(g["degenerate_sum_x_mont"])
    assert O.fp_to_mont(c[1]) == L(g["degenerate_sum_y_mont"])


def test_g1_group_law_small():
    G = O.G1_GEN
    acc = None
    for k in range(1, 12):
        acc = O.g1_add(acc, G)
        assert acc == O.g1_mul(G, k)
    assert O.g1_add(O.g1_mul(G, 4), O.g1_mul(G, 2)) == O.g1_mul(G, 6)
    P = O.g1_mul(G, 12345)
    assert O.g1_add(P, O.g1_neg(P)) is None


def test_g1_limb_layout_roundtrip():
    P = O.g1_mul(O.G1_GEN, 987654321)
    limbs = O.g1_affine_mont_limbs(P)
    assert O.g1_affine_from_mont_limbs(limbs) == P
    assert O.g1_affine_from_mont_limbs([0] * 12) is None


# ---------------------------------------------------------------- NTT (fft.rs, domain.rs)
@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5])
def test_dft_matches_bruteforce(log_n):
    rng = random.Random(log_n)
    a = [rng.randrange(O.R_MOD) for _ in range(1 << log_n)]
    w = O.root_of_unity(log_n)
    assert O.dft(a, w) == O.dft_bruteforce(a, w)
    assert O.ifft(O.best_fft(a, w), log_n) == a


def test_dft_matches_horner():
    rng = random.Random(7)
    log_n = 8
    a = [rng.randrange(O.R_MOD) for _ in range(1 << log_n)]
    w = O.root_of_unity(log_n)
    out = O.dft(a, w)
    for j in [0, 1, 5, 77, 255]:
        assert out[j] == O.poly_eval(a, pow(w, j, O.R_MOD))


def test_coset_extension_is_evaluation_on_zeta_coset():
    rng = random.Random(9)
    k, j = 4, 5
    ek = O.extended_k_for(j, k)
    assert ek == 6
    c = [rng.randrange(O.R_MOD) for _ in range(1 << k)]
    ext = O.coeff_to_extended(c, ek)
    we = O.root_of_unity(ek)
    for i in [0, 1, 2, 17, 63]:
        assert ext[i] == O.poly_eval(c, O.FR_ZETA * pow(we, i, O.R_MOD) % O.R_MOD)
    back = O.extended_to_coeff(ext, ek)
    assert back == c + [0] * ((1 << ek) - (1 << k))


def test_vanishing_division():
    k, ek = 3, 5
    t = O.t_evaluations(k, ek)
    we = O.root_of_unity(ek)
    for i in range(1 << ek):
        x = O.FR_ZETA * pow(we, i, O.R_MOD) % O.R_MOD
        assert t[i % len(t)] * (pow(x, 1 << k, O.R_MOD) - 1) % O.R_MOD == 1


# ---------------------------------------------------------------- Booth / MSM (msm.rs)
@pytest.mark.parametrize("c", list(range(1, 17)))
def test_booth_identity(c):
    rng = random.Random(c)
    nw = O.FR_BITS // c + 1
    for s in [0, 1, O.R_MOD - 1, (1 << 254) + 3] + [rng.randrange(O.R_MOD) for _ in range(30)]:
        rep = O.fr_to_bytes_le(s)
        digits = [O.get_booth_index(w, c, rep) for w in range(nw)]
        assert sum(d << (c * w) for w, d in enumerate(digits)) == s
        assert all(-(1 << (c - 1)) <= d <= (1 << (c - 1)) for d in digits)


def test_msm_booth_equals_naive():
    rng = random.Random(3)
    pts = [O.g1_mul(O.G1_GEN, rng.randrange(1, O.R_MOD)) for _ in range(6)]
    sc = [rng.randrange(O.R_MOD) for _ in range(6)]
    ref = O.msm_naive(sc, pts)
    for c in (1, 3, 4):
        assert O.msm_booth(sc, pts, c) == ref


def test_msm_srs_closed_form():
    rng = random.Random(11)
    tau = rng.randrange(O.R_MOD)
    n = 6
    bases = [O.srs_point(tau, i) for i in range(n)]
    sc = [rng.randrange(O.R_MOD) for _ in range(n)]
    assert O.msm_naive(sc, bases) == O.msm_srs_closed_form(sc, tau)


def test_window_rule():
    assert O.msm_window_size(1 << 20) == 14
    assert O.msm_window_size(1 << 16) == 12
    assert O.msm_window_size(3) == 1 and O.msm_window_size(31) == 3
