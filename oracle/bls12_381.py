"""Oracle: pure-Python big-integer restatement of the BLS12-381 arithmetic on the
prover's MSM / NTT path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``.  Product code never imports it, and no product result is ever
computed by it.

Naming: ``fr`` is the 255-bit scalar field (the reference calls it ``Fq``,
curves/src/bls12_381/fq.rs:1-3,29); ``fp`` is the 381-bit base field (the
reference's ``Fp``, curves/src/bls12_381/fp.rs:1-3,98).

What pins each part (tests/test_oracle_python.py):
  * field constants and Montgomery encodings -- the byte/limb KATs of
    fq.rs:934-1230 and fp.rs:1258-1290 (copied into tests/golden/reference_kats.json);
  * G1 group law -- KATs g1.rs:1153-1186 (2G) and g1.rs:1266-1298 (degenerate add),
    the generator of curves/README.md, on-curve and order-r checks;
  * NTT -- brute-force DFT on tiny sizes, inverse(forward) = id, and Horner
    evaluation at the domain points (independent closed form);
  * Booth digits -- the identity sum_w d_w 2^(cw) = s (msm.rs:547-590);
  * MSM -- brute force (sum of scalar multiples), and for SRS bases the closed
    form [sum_i s_i tau^i] G.
Nothing here is "parity unpinned".
"""
from __future__ import annotations

import hashlib

# --------------------------------------------------------------------------
# Field constants (fq.rs:43-126, fp.rs:20-57)
# --------------------------------------------------------------------------
R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P_MOD = int(
    "1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f624"
    "1eabfffeb153ffffb9feffffffffaaab", 16)
FR_BITS = 255
FR_S = 32                                # fq.rs:126
FR_GENERATOR = 7                         # fq.rs:31-40
FR_R = pow(2, 256, R_MOD)                # Montgomery radix for fr (fq.rs:84-97)
FP_R = pow(2, 384, P_MOD)                # Montgomery radix for fp (fp.rs:55-66)
FR_RINV = pow(2, -256, R_MOD)
FP_RINV = pow(2, -384, P_MOD)
FR_ROOT_OF_UNITY = pow(FR_GENERATOR, (R_MOD - 1) >> FR_S, R_MOD)   # fq.rs:70-77
FR_ZETA = 0xAC45A4010001A40200000000FFFFFFFF                          # fq.rs:700-710 (canonical)
FR_DELTA = pow(FR_GENERATOR, 1 << FR_S, R_MOD)                     # fq.rs:501-509
FR_TWO_INV = (R_MOD + 1) // 2                                       # fq.rs:480-487

G1_B = 4                                 # g1.rs:40-50
G1_GEN = (
    3685416753713387016781088315183077757961620795782546409894578378688607592378376318836054947676345821548104185464507,
    1339506544944476473020471379941921221584933875938349620426543736416511423956333506472724655353366534992391756441569,
)


# --------------------------------------------------------------------------
# Montgomery encodings (limb layout of blst_fr / blst_fp, fq.rs:27-29, fp.rs:94-98)
# --------------------------------------------------------------------------
def to_limbs(x: int, n: int, bits: int = 64) -> list[int]:
    mask = (1 << bits) - 1
    return [(x >> (bits * i)) & mask for i in range(n)]


def from_limbs(limbs, bits: int = 64) -> int:
    v = 0
    for i, l in enumerate(limbs):
        v |= int(l) << (bits * i)
    return v


def fr_to_mont(x: int) -> int:
    return (x % R_MOD) * FR_R % R_MOD


def fr_from_mont(m: int) -> int:
    return m * FR_RINV % R_MOD


def fp_to_mont(x: int) -> int:
    return (x % P_MOD) * FP_R % P_MOD


def fp_from_mont(m: int) -> int:
    return m * FP_RINV % P_MOD


def fr_mont_mul(a_m: int, b_m: int) -> int:
    """blst_fr_mul on Montgomery representatives: a*b*2^-256 mod r."""
    return a_m * b_m * FR_RINV % R_MOD


def fp_mont_mul(a_m: int, b_m: int) -> int:
    return a_m * b_m * FP_RINV % P_MOD


def fr_to_bytes_le(x: int) -> bytes:
    """Canonical 32-byte little-endian encoding (fq.rs:532, to_repr/to_bytes_le)."""
    return (x % R_MOD).to_bytes(32, "little")


def fr_from_bytes_le(b: bytes):
    """fq.rs from_bytes_le: None when the integer is >= r."""
    v = int.from_bytes(b, "little")
    return v if v < R_MOD else None


def fr_from_uniform_bytes(b: bytes) -> int:
    """fq.rs:712-736: a0*R2 + a1*R3 in Montgomery form, i.e. the 512-bit LE
    integer reduced mod r (returned canonical)."""
    assert len(b) == 64
    return int.from_bytes(b, "little") % R_MOD


# --------------------------------------------------------------------------
# Seeded synthetic streams shared by the product's device generator and the oracle.
# splitmix64 (Steele, Lea, Flood 2014) keyed by (seed, stream, index).
# --------------------------------------------------------------------------
M64 = (1 << 64) - 1
SM_GAMMA = 0x9E3779B97F4A7C15


def splitmix64(x: int) -> int:
    x = (x + SM_GAMMA) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def stream_key(seed: int, stream: int) -> int:
    return splitmix64((seed ^ ((stream * 0xD1B54A32D192ED03) & M64)) & M64)


def fr_from_seed(seed: int, stream: int, index: int) -> int:
    """Element `index` of stream `stream`: eight splitmix64 words form 64
    uniform bytes, reduced as Fq::from_uniform_bytes does (fq.rs:712-736)."""
    key = stream_key(seed, stream)
    words = [splitmix64((key + index * 8 + j) & M64) for j in range(8)]
    return fr_from_uniform_bytes(b"".join(w.to_bytes(8, "little") for w in words))


def fr_vector_from_seed(seed: int, stream: int, n: int, offset: int = 0) -> list[int]:
    return [fr_from_seed(seed, stream, offset + i) for i in range(n)]


# --------------------------------------------------------------------------
# G1 on y^2 = x^3 + 4 over fp (g1.rs).  Affine points are (x, y) canonical ints,
# the identity is None.
# --------------------------------------------------------------------------
def g1_is_on_curve(P) -> bool:
    if P is None:
        return True
    x, y = P
    return (y * y - x * x * x - G1_B) % P_MOD == 0


def g1_neg(P):
    if P is None:
        return None
    return (P[0], (-P[1]) % P_MOD)


def g1_add(P, Q):
    """Textbook affine chord-and-tangent law (includes doubling and P + (-P))."""
    if P is None:
        return Q
    if Q is None:
        return P
    x1, y1 = P
    x2, y2 = Q
    if x1 == x2:
        if (y1 + y2) % P_MOD == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P_MOD) % P_MOD
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P_MOD) % P_MOD
    x3 = (lam * lam - x1 - x2) % P_MOD
    y3 = (lam * (x1 - x3) - y1) % P_MOD
    return (x3, y3)


def g1_double(P):
    return g1_add(P, P)


# Jacobian arithmetic (x = X/Z^2, y = Y/Z^3; g1.rs:1052-1053) keeps scalar
# multiplication fast enough for the closed-form checks.
def _jac_double(X, Y, Z):
    if Z == 0 or Y == 0:
        return (1, 1, 0)
    A = X * X % P_MOD
    B = Y * Y % P_MOD
    C = B * B % P_MOD
    D = 2 * ((X + B) ** 2 - A - C) % P_MOD
    E = 3 * A % P_MOD
    F = E * E % P_MOD
    X3 = (F - 2 * D) % P_MOD
    Y3 = (E * (D - X3) - 8 * C) % P_MOD
    Z3 = 2 * Y * Z % P_MOD
    return (X3, Y3, Z3)


def _jac_add(P1, P2):
    X1, Y1, Z1 = P1
    X2, Y2, Z2 = P2
    if Z1 == 0:
        return P2
    if Z2 == 0:
        return P1
    Z1Z1 = Z1 * Z1 % P_MOD
    Z2Z2 = Z2 * Z2 % P_MOD
    U1 = X1 * Z2Z2 % P_MOD
    U2 = X2 * Z1Z1 % P_MOD
    S1 = Y1 * Z2 * Z2Z2 % P_MOD
    S2 = Y2 * Z1 * Z1Z1 % P_MOD
    if U1 == U2:
        if S1 == S2:
            return _jac_double(X1, Y1, Z1)
        return (1, 1, 0)
    H = (U2 - U1) % P_MOD
    I = (2 * H) ** 2 % P_MOD
    J = H * I % P_MOD
    rr = 2 * (S2 - S1) % P_MOD
    V = U1 * I % P_MOD
    X3 = (rr * rr - J - 2 * V) % P_MOD
    Y3 = (rr * (V - X3) - 2 * S1 * J) % P_MOD
    Z3 = ((Z1 + Z2) ** 2 - Z1Z1 - Z2Z2) * H % P_MOD
    return (X3, Y3, Z3)


def jac_from_affine(P):
    return (1, 1, 0) if P is None else (P[0], P[1], 1)


def jac_to_affine(J):
    X, Y, Z = J
    if Z % P_MOD == 0:
        return None
    zi = pow(Z, -1, P_MOD)
    zi2 = zi * zi % P_MOD
    return (X * zi2 % P_MOD, Y * zi2 * zi % P_MOD)


def g1_mul(P, k: int):
    """[k]P by left-to-right double-and-add (k taken as a non-negative int)."""
    if P is None or k == 0:
        return None
    acc = (1, 1, 0)
    base = jac_from_affine(P)
    for bit in bin(k)[2:]:
        acc = _jac_double(*acc)
        if bit == "1":
            acc = _jac_add(acc, base)
    return jac_to_affine(acc)


def g1_sum(points):
    acc = (1, 1, 0)
    for P in points:
        acc = _jac_add(acc, jac_from_affine(P))
    return jac_to_affine(acc)


def g1_to_compressed(P) -> bytes:
    """ZCash BLS12-381 48-byte compressed encoding (g1.rs:326-334 via blst):
    big-endian x, flag 0x80 = compressed, 0x40 = infinity, 0x20 = y is the
    lexicographically larger root."""
    if P is None:
        out = bytearray(48)
        out[0] = 0xC0
        return bytes(out)
    x, y = P
    out = bytearray(x.to_bytes(48, "big"))
    out[0] |= 0x80
    if y > (P_MOD - 1) // 2:
        out[0] |= 0x20
    return bytes(out)


def g1_to_uncompressed(P) -> bytes:
    """96-byte big-endian x || y, flag 0x40 for the identity (g1.rs:336-344)."""
    if P is None:
        out = bytearray(96)
        out[0] = 0x40
        return bytes(out)
    return P[0].to_bytes(48, "big") + P[1].to_bytes(48, "big")


class PointDecodeError(ValueError):
    """reason: 'encoding', 'curve' or 'subgroup'."""

    def __init__(self, reason: str):
        super().__init__(reason)
        self.reason = reason


def g1_in_subgroup(P) -> bool:
    """[r] P = O: the test that defines G1 (r prime, cofactor prime to r)."""
    return P is None or g1_mul(P, R_MOD) is None


def g1_from_compressed(b: bytes, check_subgroup: bool = True):
    """Inverse of g1_to_compressed per the ZCash serialisation (from_compressed,
    g1.rs:367-384 via blst): flags checked, x < p, y = sqrt(x^3 + 4) chosen by the
    0x20 flag; raises PointDecodeError."""
    if len(b) != 48:
        raise PointDecodeError("encoding")
    flags = b[0] >> 5
    x = int.from_bytes(bytes([b[0] & 0x1F]) + b[1:], "big")
    if not flags & 4:
        raise PointDecodeError("encoding")
    if flags & 2:
        if flags & 1 or x:
            raise PointDecodeError("encoding")
        return None
    if x >= P_MOD:
        raise PointDecodeError("encoding")
    rhs = (x * x * x + G1_B) % P_MOD
    y = pow(rhs, (P_MOD + 1) // 4, P_MOD)   # p = 3 mod 4
    if y * y % P_MOD != rhs:
        raise PointDecodeError("curve")
    if (y > (P_MOD - 1) // 2) != bool(flags & 1):
        y = P_MOD - y
    if check_subgroup and not g1_in_subgroup((x, y)):
        raise PointDecodeError("subgroup")
    return (x, y)


def g1_from_uncompressed(b: bytes, check_subgroup: bool = True):
    """Inverse of g1_to_uncompressed (from_uncompressed / _unchecked, g1.rs:337-365):
    flags, x, y < p, the curve equation, and membership of G1 when asked."""
    if len(b) != 96:
        raise PointDecodeError("encoding")
    flags = b[0] >> 5
    x = int.from_bytes(bytes([b[0] & 0x1F]) + b[1:48], "big")
    y = int.from_bytes(b[48:], "big")
    if flags & 4 or flags & 1:
        raise PointDecodeError("encoding")
    if flags & 2:
        if x or y:
            raise PointDecodeError("encoding")
        return None
    if x >= P_MOD or y >= P_MOD:
        raise PointDecodeError("encoding")
    if not g1_is_on_curve((x, y)):
        raise PointDecodeError("curve")
    if check_subgroup and not g1_in_subgroup((x, y)):
        raise PointDecodeError("subgroup")
    return (x, y)


def g1_affine_mont_limbs(P) -> list[int]:
    """blst_p1_affine layout: x[6] || y[6] u64 LE Montgomery; identity = zeros
    (g1.rs:32-34)."""
    if P is None:
        return [0] * 12
    return to_limbs(fp_to_mont(P[0]), 6) + to_limbs(fp_to_mont(P[1]), 6)


def g1_affine_from_mont_limbs(limbs):
    limbs = [int(v) for v in limbs]
    if all(v == 0 for v in limbs):
        return None
    return (fp_from_mont(from_limbs(limbs[:6])), fp_from_mont(from_limbs(limbs[6:12])))


def g1_jacobian_from_mont_limbs(limbs):
    """blst_p1 layout X[6] || Y[6] || Z[6] Montgomery, Jacobian (g1.rs:493-495)."""
    limbs = [int(v) for v in limbs]
    X = fp_from_mont(from_limbs(limbs[0:6]))
    Y = fp_from_mont(from_limbs(limbs[6:12]))
    Z = fp_from_mont(from_limbs(limbs[12:18]))
    return jac_to_affine((X, Y, Z))


# --------------------------------------------------------------------------
# Booth digits (restates msm.rs:13-55, get_booth_index)
# --------------------------------------------------------------------------
def get_booth_index(window_index: int, window_size: int, el: bytes) -> int:
    skip_bits = max(window_index * window_size - 1, 0)
    skip_bytes = skip_bits // 8
    v = bytearray(4)
    for i in range(4):
        if skip_bytes + i < len(el):
            v[i] = el[skip_bytes + i]
    tmp = int.from_bytes(bytes(v), "little")
    if window_index == 0:
        tmp = (tmp << 1) & 0xFFFFFFFF
    tmp >>= skip_bits - skip_bytes * 8
    tmp &= (1 << (window_size + 1)) - 1
    sign = (tmp & (1 << window_size)) == 0
    tmp = (tmp + 1) >> 1
    if sign:
        return tmp
    return -((~(tmp - 1)) & ((1 << window_size) - 1))


def msm_window_size(n: int) -> int:
    """Window rule of msm_serial / msm_best (msm.rs:343-349, 464-470)."""
    import math
    if n < 4:
        return 1
    if n < 32:
        return 3
    return int(math.ceil(math.log(float(n))))


# --------------------------------------------------------------------------
# MSM references
# --------------------------------------------------------------------------
def msm_naive(scalars, points):
    """sum_i [s_i] P_i -- the definition msm_best/msm_parallel must equal
    (msm.rs:592-621, g1.rs:1582-1600)."""
    acc = (1, 1, 0)
    for s, P in zip(scalars, points):
        if P is None or s % R_MOD == 0:
            continue
        acc = _jac_add(acc, jac_from_affine(g1_mul(P, s % R_MOD)))
    return jac_to_affine(acc)


def msm_booth(scalars, points, c: int):
    """Windowed signed-Booth evaluation exactly as msm_serial orders it
    (msm.rs:338-423), without the bucket tricks: sum_w 2^(cw) sum_i d_{i,w} P_i."""
    nw = FR_BITS // c + 1
    reprs = [fr_to_bytes_le(s) for s in scalars]
    acc = (1, 1, 0)
    for w in reversed(range(nw)):
        for _ in range(c):
            acc = _jac_double(*acc)
        for rep, P in zip(reprs, points):
            d = get_booth_index(w, c, rep)
            if d > 0:
                acc = _jac_add(acc, jac_from_affine(g1_mul(P, d)))
            elif d < 0:
                acc = _jac_add(acc, jac_from_affine(g1_neg(g1_mul(P, -d))))
    return jac_to_affine(acc)


def poly_eval(coeffs, x: int) -> int:
    """Horner evaluation mod r (proofs/src/utils/arithmetic.rs:59-87)."""
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R_MOD
    return acc


def srs_point(tau: int, i: int):
    """g_i = [tau^i] G (params.rs:154-172)."""
    return g1_mul(G1_GEN, pow(tau, i, R_MOD))


def msm_srs_closed_form(scalars, tau: int):
    """With bases g_i = [tau^i]G, sum_i s_i g_i = [poly_s(tau)] G."""
    return g1_mul(G1_GEN, poly_eval(list(scalars), tau))


# --------------------------------------------------------------------------
# NTT / EvaluationDomain (curves/src/fft.rs, proofs/src/poly/domain.rs)
# --------------------------------------------------------------------------
def root_of_unity(log_n: int) -> int:
    """omega = ROOT^(2^(S - log_n)) (domain.rs:72-91)."""
    return pow(FR_ROOT_OF_UNITY, 1 << (FR_S - log_n), R_MOD)


def compute_twiddles(omega: int, log_n: int) -> list[int]:
    """fft.rs:31-40: [1, w, ..., w^(n/2-1)]."""
    tw, w = [], 1
    for _ in range(1 << (log_n - 1)):
        tw.append(w)
        w = w * omega % R_MOD
    return tw


def dft(a, omega: int):
    """A_j = sum_i a_i omega^(ij) by recursive radix-2 (any power-of-two size)."""
    n = len(a)
    if n == 1:
        return [a[0] % R_MOD]
    even = dft(a[0::2], omega * omega % R_MOD)
    odd = dft(a[1::2], omega * omega % R_MOD)
    out = [0] * n
    w = 1
    for k in range(n // 2):
        t = w * odd[k] % R_MOD
        out[k] = (even[k] + t) % R_MOD
        out[k + n // 2] = (even[k] - t) % R_MOD
        w = w * omega % R_MOD
    return out


def dft_bruteforce(a, omega: int):
    n = len(a)
    return [sum(a[i] * pow(omega, i * j, R_MOD) for i in range(n)) % R_MOD for j in range(n)]


def best_fft(a, omega: int):
    """best_fft (fft.rs:49-101): in-place DFT with omega, natural order."""
    return dft(list(a), omega)


def ifft(a, log_n: int):
    """domain.rs:389-396: DFT with omega^-1 then x 1/n."""
    w_inv = pow(root_of_unity(log_n), -1, R_MOD)
    ninv = pow(1 << log_n, -1, R_MOD)
    return [v * ninv % R_MOD for v in dft(list(a), w_inv)]


def distribute_powers_zeta(a, into_coset: bool):
    """domain.rs:398-414: a_i *= [1, g, g^2][i mod 3] with g = zeta (into) or
    zeta^2 (out of the coset)."""
    z = FR_ZETA
    z2 = z * z % R_MOD
    powers = (1, z, z2) if into_coset else (1, z2, z)
    return [v * powers[i % 3] % R_MOD for i, v in enumerate(a)]


def extended_k_for(j: int, k: int) -> int:
    """EvaluationDomain::new (domain.rs:54-70): smallest ext_k with 2^ext_k >= n(j-1)."""
    n = 1 << k
    ek = k
    while (1 << ek) < n * (j - 1):
        ek += 1
    return ek


def coeff_to_extended(coeffs, ext_k: int):
    """domain.rs:285-305 + fft.rs:110-186: zeta-coset, zero pad, DFT on 2^ext_k.
    Equals evaluation of the polynomial at zeta * omega_ext^j."""
    a = distribute_powers_zeta(coeffs, True) + [0] * ((1 << ext_k) - len(coeffs))
    return dft(a, root_of_unity(ext_k))


def extended_to_coeff(evals, ext_k: int):
    """domain.rs:312-328: iDFT on 2^ext_k, x 1/2^ext_k, then zeta^-(i mod 3)."""
    return distribute_powers_zeta(ifft(evals, ext_k), False)


def t_evaluations(k: int, ext_k: int):
    """domain.rs:104-140: 1/((zeta w_ext^i)^n - 1) for i < 2^(ext_k-k)."""
    n = 1 << k
    we = root_of_unity(ext_k)
    out = []
    for i in range(1 << (ext_k - k)):
        v = pow(FR_ZETA * pow(we, i, R_MOD) % R_MOD, n, R_MOD)
        out.append(pow((v - 1) % R_MOD, -1, R_MOD))
    return out


def divide_by_vanishing_poly(evals, k: int, ext_k: int):
    """domain.rs:361-380."""
    t = t_evaluations(k, ext_k)
    m = len(t) - 1
    return [v * t[i & m] % R_MOD for i, v in enumerate(evals)]


def sha256_fr_vector(values) -> str:
    """Digest of a vector of canonical fr values as 32-byte LE words."""
    h = hashlib.sha256()
    for v in values:
        h.update(fr_to_bytes_le(v))
    return h.hexdigest()
