"""PLONK/KZG prover for the synthetic five-column circuit, over a pluggable backend.

The proof follows midnight-proofs' `create_proof` (proofs/src/plonk/prover.rs:432-474
-> compute_trace :66-299 -> finalise_proof :302-423) with the KZG multi-open of
proofs/src/poly/kzg/mod.rs:110-289, for a circuit with

  advice a, b, c, d, e (all in the permutation, in that order), no instances,
  fixed q_m, q_l, q_r, q_o, q_c, one gate  a*b*q_m + a*q_l + b*q_r - c*q_o + q_c,
  minimum degree 5  =>  cs degree 5, quotient degree 4, extended k = k + 2,
  permutation chunks of 3 columns  =>  two grand-product sets,
  blinding factors 6 (circuit.rs:2203-2246).

The transcript is CircuitTranscript<Blake2bState> (transcript/mod.rs:82-138,
implementors.rs:13-29, 138-185): Blake2b-512 keyed "Domain separator for transcript";
absorb = 0x01 || bytes; squeeze = 0x00 then the (non-destructive) digest, read as a
512-bit little-endian integer mod r.  G1 goes into the proof compressed (48 B),
scalars as 32 B little-endian canonical.

All polynomial arithmetic runs through a backend object.  `GpuBackend` (this file)
is the product: device vectors, HIP kernels for every field operation, resident SRS
bases for the MSMs.  A second backend with the same methods can run the identical
sequence elsewhere; since every operation is exact, equal inputs give byte-identical
proofs.

Where this differs from the reference, by design:
  * the verifying-key hash (plonk/mod.rs:249-283) hashes the VK's commitments as the
    reference does, but the `Debug` text of the pinned domain and constraint system
    is replaced by a fixed descriptor string;
  * the prover's random draws (advice blinding rows, permutation blinding rows,
    quotient-limb blinders, prover.rs:607-616, 104-110, 774-791) come from a seeded
    scalar stream instead of the caller's RngCore / OsRng, so proofs are reproducible;
  * the permutation commitments are taken in the Lagrange basis; the reference uses
    the LagrangeDelta basis (permutation/prover.rs:236-245), which gives the same
    group element;
  * the per-query evaluations multi_open computes for its intermediate sets
    (kzg/mod.rs:179-188) are never read by the prover and are not recomputed.
"""
from __future__ import annotations

import hashlib
import time

import numpy as np

# ---------------------------------------------------------------- field constants
R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P_MOD = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
FR_R = (1 << 256) % R_MOD
FR_RINV = pow(1 << 256, -1, R_MOD)
FP_RINV = pow(1 << 384, -1, P_MOD)
GENERATOR = 7                                           # fq.rs:31-40
ROOT_OF_UNITY = pow(GENERATOR, (R_MOD - 1) >> 32, R_MOD)  # fq.rs:70-77
ZETA = pow(GENERATOR, (R_MOD - 1) // 3, R_MOD)           # fq.rs:700-710 (coset shift)
DELTA = pow(GENERATOR, 1 << 32, R_MOD)                   # fq.rs:501-509
M64 = (1 << 64) - 1

NUM_ADVICE = 5
NUM_FIXED = 5
BLINDING_FACTORS = 6            # circuit.rs:2203-2246 for this circuit
CS_DEGREE = 5                   # set_minimum_degree(5)
QUOTIENT_DEGREE = CS_DEGREE - 1
CHUNK_LEN = CS_DEGREE - 2       # permutation/prover.rs:102
RNG_STREAM = 0x524E47           # seeded stream of the prover's random draws
MAX_CYCLES = 1 << 14


def mont_limbs(v: int) -> np.ndarray:
    """Canonical integer -> 4 u64 Montgomery limbs (Fq's in-memory form)."""
    m = (v % R_MOD) * (1 << 256) % R_MOD
    return np.array([(m >> (64 * i)) & M64 for i in range(4)], dtype=np.uint64)


def from_mont_limbs(limbs) -> int:
    v = 0
    for i, w in enumerate(np.asarray(limbs, dtype=np.uint64).reshape(4)):
        v |= int(w) << (64 * i)
    return v * FR_RINV % R_MOD


def fr_repr(v: int) -> bytes:
    """Fq::to_repr: canonical 32 bytes little-endian (fq.rs:309-343)."""
    return (v % R_MOD).to_bytes(32, "little")


def omega_for(k: int) -> int:
    return pow(ROOT_OF_UNITY, 1 << (32 - k), R_MOD)


def g1_compress(aff12) -> bytes:
    """G1Affine::to_compressed (ZCash format): x big-endian with flag bits 0x80
    (compressed), 0x40 (infinity), 0x20 (y is the larger root)."""
    limbs = [int(w) for w in np.asarray(aff12, dtype=np.uint64).reshape(12)]
    if not any(limbs):
        return bytes([0xC0]) + bytes(47)
    xm = sum(limbs[i] << (64 * i) for i in range(6))
    ym = sum(limbs[6 + i] << (64 * i) for i in range(6))
    x = xm * FP_RINV % P_MOD
    y = ym * FP_RINV % P_MOD
    out = bytearray(x.to_bytes(48, "big"))
    out[0] |= 0x80
    if y > (P_MOD - 1) // 2:
        out[0] |= 0x20
    return bytes(out)


def g1_raw_bytes(aff12) -> bytes:
    """SerdeFormat::RawBytesUnchecked: the 96-byte Montgomery affine limbs."""
    return np.asarray(aff12, dtype=np.uint64).reshape(12).astype("<u8").tobytes()


class Transcript:
    """CircuitTranscript<Blake2bState> (transcript/mod.rs:82-138)."""

    def __init__(self):
        self.h = hashlib.blake2b(digest_size=64, key=b"Domain separator for transcript")
        self.proof = bytearray()

    def common_fr(self, v: int):
        self.h.update(b"\x01" + fr_repr(v))

    def write_fr(self, v: int):
        self.common_fr(v)
        self.proof += fr_repr(v)

    def write_point(self, compressed: bytes):
        self.h.update(b"\x01" + compressed)
        self.proof += compressed

    def squeeze(self) -> int:
        self.h.update(b"\x00")
        return int.from_bytes(self.h.digest(), "little") % R_MOD


# ---------------------------------------------------------------- seeded structure
def _splitmix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def _stream(seed: int, tag: int, n: int) -> np.ndarray:
    key = _splitmix(np.array([(seed ^ (tag * 0xD1B54A32D192ED03)) & M64], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        return _splitmix(np.arange(n, dtype=np.uint64) + key)


class SyntheticCircuit:
    """The synthetic circuit's layout and witness recipe at size 2^k.

    Usable rows u = n - 7 (prover.rs:575).  Row i < u is a mul gate (q_m = q_o = 1,
    c = a*b) or an add gate (q_l = q_r = q_o = 1, c = a + b), chosen by bit 0 of a
    seeded splitmix64 stream; rows >= u have all selectors zero.  Copy constraints
    are min(2^14, 4u/16) cycles of length 2..8 over distinct cells of columns a, b, d,
    e (rows < u): every cell of a cycle holds the value of the cycle's first cell, and
    the permutation maps each cycle cell to the next one (the last to the first).
    Witness values are elements of the seeded Fr stream (seed, column); c is derived
    after the copies, so gates and copies both hold.
    """

    def __init__(self, k: int, seed: int):
        if not 4 <= k <= 26:
            raise ValueError("k out of range")
        self.k = k
        self.n = n = 1 << k
        self.seed = seed
        self.usable = u = n - BLINDING_FACTORS - 1
        self.is_mul = (_stream(seed, 1, u) & np.uint64(1)).astype(bool)
        # cycles over cells of a, b, d, e: flat cell id = slot * u + row, slot 0..3
        ncyc = min(MAX_CYCLES, (4 * u) // 16)
        lens = 2 + (_stream(seed, 2, ncyc) % np.uint64(7)).astype(np.int64)
        order = np.argsort(_stream(seed, 3, 4 * u), kind="stable")
        total = int(lens.sum())
        cells = order[:total]
        starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
        cyc_of = np.repeat(np.arange(ncyc), lens)
        first = cells[starts[cyc_of]]                # first cell of each cell's cycle
        nxt_pos = np.arange(total) + 1
        wrap = nxt_pos == (starts[cyc_of] + lens[cyc_of])
        nxt_pos[wrap] = starts[cyc_of][wrap]
        self.copy_cells = cells                      # cells overwritten by the copy
        self.copy_src = first
        slot_col = np.array([0, 1, 3, 4], dtype=np.int64)   # a, b, d, e in permutation order

        def perm_index(cell):
            return slot_col[cell // u] * n + cell % u
        mapping = np.arange(NUM_ADVICE * n, dtype=np.int64)
        mapping[perm_index(cells)] = perm_index(cells[nxt_pos])
        self.mapping = mapping                       # cell (col, row) -> col' * n + row'
        self.num_cycles = ncyc

    def fixed_rows(self):
        """Row indices set to one in q_m, q_l, q_r, q_o, q_c."""
        u = self.usable
        mul = np.nonzero(self.is_mul)[0]
        add = np.nonzero(~self.is_mul)[0]
        allr = np.arange(u)
        return [mul, add, add, allr, np.zeros(0, dtype=np.int64)]

    def witness(self, be, wseed: int | None = None):
        """Advice values a..e (length n; rows >= u zero, the prover blinds them).  The
        values come from the seeded stream of wseed (default: the layout seed)."""
        n, u = self.n, self.usable
        ws = self.seed if wseed is None else wseed
        cols = [be.from_seed(ws, 16 + s, u) for s in range(4)]   # a, b, d, e
        flat = be.cat(cols)
        flat = be.scatter(flat, self.copy_cells, be.gather(flat, self.copy_src))
        a, b, d, e = (be.slice(flat, s * u, (s + 1) * u) for s in range(4))
        c = be.select(self.is_mul, be.mul(a, b), be.add(a, b))
        pad = be.zeros(n - u)
        return [be.cat([v, pad]) for v in (a, b, c, d, e)]


# ---------------------------------------------------------------- keys
class Params:
    """ParamsKZG from unsafe_setup (params.rs:154-210): monomial bases g_i = [tau^i]G,
    Lagrange bases [L_i(tau)]G, L_i(tau) = (tau^n - 1)/n * omega^i / (tau - omega^i), and
    their suffix sums g_lagrange_delta (params.rs:47-57, 186), the bases of the permutation
    commitments."""

    def __init__(self, be, k: int, tau: int, g=None):
        """g: monomial bases of the same tau with at least 2^k points to share (commits
        use the prefix g[..len], kzg/mod.rs:99-103); generated when absent."""
        self.k = k
        self.n = n = 1 << k
        self.tau = tau % R_MOD
        if g is not None and g.n < n:
            raise ValueError("shared monomial bases are shorter than 2^k")
        self.g = g if g is not None else be.srs_monomial(self.tau, n)
        w = be.powers(omega_for(k), n)
        den = be.batch_invert(be.add_scalar(be.neg(w), self.tau))
        c = (pow(self.tau, n, R_MOD) - 1) * pow(n, -1, R_MOD) % R_MOD
        self.g_lagrange = be.srs_from_scalars(be.scale(be.mul(w, den), c))
        self.g_lagrange_delta = be.srs_suffix_sum(self.g_lagrange)


class ProvingKey:
    pass


def keygen(be, params: Params, circuit: SyntheticCircuit) -> ProvingKey:
    """keygen_vk / keygen_pk (plonk/keygen.rs:240-403) for the synthetic circuit."""
    k, n = circuit.k, circuit.n
    if params.k != k:
        raise ValueError("params and circuit sizes differ")
    dom = be.domain(CS_DEGREE, k)
    pk = ProvingKey()
    pk.k, pk.n, pk.dom = k, n, dom
    pk.ext_n = 1 << dom.extended_k
    one = be.const(1, 1)
    fixed_values = []
    for rows in circuit.fixed_rows():
        v = be.zeros(n)
        if len(rows):
            v = be.scatter(v, rows, be.broadcast(one, len(rows)))
        fixed_values.append(v)
    # permutation: sigma_j(omega^i) = delta^j' omega^i' for (j', i') = mapping(j, i)
    w = be.powers(omega_for(k), n)
    deltaomega = be.cat([be.scale(w, pow(DELTA, j, R_MOD)) for j in range(NUM_ADVICE)])
    sig_all = be.gather(deltaomega, circuit.mapping)
    sigma_values = [be.slice(sig_all, j * n, (j + 1) * n) for j in range(NUM_ADVICE)]
    coms = be.commit_many(params, [(v, "lagrange") for v in fixed_values + sigma_values])
    pk.fixed_commitments = coms[:NUM_FIXED]
    pk.permutation_commitments = coms[NUM_FIXED:]
    pk.fixed_values = fixed_values
    pk.fixed_polys = [dom.lagrange_to_coeff(v) for v in fixed_values]
    pk.fixed_cosets = [dom.coeff_to_extended(p) for p in pk.fixed_polys]
    pk.sigma_values = sigma_values
    pk.sigma_polys = [dom.lagrange_to_coeff(v) for v in sigma_values]
    pk.sigma_cosets = [dom.coeff_to_extended(p) for p in pk.sigma_polys]
    # l0, l_last, l_active_row (keygen.rs:366-403)
    def lagrange_unit(rows):
        v = be.zeros(n)
        return be.scatter(v, np.asarray(rows, dtype=np.int64), be.broadcast(one, len(rows)))
    pk.l0 = dom.coeff_to_extended(dom.lagrange_to_coeff(lagrange_unit([0])))
    l_blind = dom.coeff_to_extended(dom.lagrange_to_coeff(
        lagrange_unit(list(range(n - BLINDING_FACTORS, n)))))
    pk.l_last = dom.coeff_to_extended(dom.lagrange_to_coeff(lagrange_unit([n - BLINDING_FACTORS - 1])))
    pk.l_active_row = be.add_scalar(be.neg(be.add(pk.l_last, l_blind)), 1)
    pk.omega_pows = be.powers(omega_for(k), n)
    # the coset points zeta * omega_ext^i of the extended domain
    pk.coset_x = be.scale(be.powers(omega_for(dom.extended_k), pk.ext_n), ZETA)
    pk.transcript_repr = vk_transcript_repr(k, dom.extended_k, pk.fixed_commitments,
                                            pk.permutation_commitments)
    return pk


def vk_transcript_repr(k: int, ext_k: int, fixed_coms, perm_coms) -> int:
    """VerifyingKey::from_parts hash (plonk/mod.rs:249-283): Blake2b-512 personal
    "Halo2-Verify-Key" over version 0x03, k, the commitments (96-byte raw form), and a
    descriptor of the domain and constraint system; read with from_uniform_bytes."""
    h = hashlib.blake2b(digest_size=64, person=b"Halo2-Verify-Key")
    buf = bytearray([0x03, k])
    buf += len(fixed_coms).to_bytes(4, "little")
    for c in fixed_coms:
        buf += g1_raw_bytes(c)
    buf += len(perm_coms).to_bytes(4, "little")
    for c in perm_coms:
        buf += g1_raw_bytes(c)
    buf += (f"PinnedEvaluationDomain {{ k: {k}, extended_k: {ext_k}, omega: 0x{omega_for(k):064x} }}").encode()
    buf += (b"SyntheticPlonk { advice: [a, b, c, d, e], fixed: [q_m, q_l, q_r, q_o, q_c], "
            b"gate: a*b*q_m + a*q_l + b*q_r - c*q_o + q_c, permutation: [a, b, c, d, e], "
            b"minimum_degree: 5 }")
    h.update(bytes(buf))
    return int.from_bytes(h.digest(), "little") % R_MOD


# ---------------------------------------------------------------- prover
def _rotate(x: int, rot: int, k: int) -> int:
    """EvaluationDomain::rotate_omega (domain.rs)."""
    w = omega_for(k)
    return x * pow(w if rot >= 0 else pow(w, -1, R_MOD), abs(rot), R_MOD) % R_MOD


def _powers(x: int, m: int):
    out, t = [], 1
    for _ in range(m):
        out.append(t)
        t = t * x % R_MOD
    return out


def numerator_composed(be, pk: "ProvingKey", adv_c, z_c, y: int, beta: int, gamma: int):
    """evaluate_numerator (evaluation.rs:818-940) for this circuit from element-wise
    vector operations: the gate, then the permutation identities in the reference's
    order, each step v <- v y + term."""
    bf = BLINDING_FACTORS
    scale = pk.ext_n // pk.n
    sets = [list(range(0, CHUNK_LEN)), list(range(CHUNK_LEN, NUM_ADVICE))]
    qm, ql, qr, qo, qc = pk.fixed_cosets
    a, b, c = adv_c[0], adv_c[1], adv_c[2]
    val = be.add(be.add(be.add(be.mul(be.mul(a, b), qm), be.mul(a, ql)), be.mul(b, qr)),
                 be.sub(qc, be.mul(c, qo)))
    rot_next = scale
    rot_last = -(bf + 1) * scale
    # l0 (1 - z_0)
    val = be.add(be.scale(val, y), be.mul(be.add_scalar(be.neg(z_c[0]), 1), pk.l0))
    # l_last (z_l^2 - z_l)
    zl = z_c[-1]
    val = be.add(be.scale(val, y), be.mul(be.sub(be.mul(zl, zl), zl), pk.l_last))
    # l0 (z_i - z_(i-1)(omega^last X))
    for s in range(1, len(z_c)):
        val = be.add(be.scale(val, y), be.mul(be.sub(z_c[s], be.roll(z_c[s - 1], rot_last)), pk.l0))
    # l_active (z(omega X) prod(p + beta s + gamma) - z(X) prod(p + delta^j beta X + gamma))
    for s, cols in enumerate(sets):
        left = be.roll(z_c[s], rot_next)
        right = z_c[s]
        for j in cols:
            left = be.mul(left, be.add_scalar(be.add(adv_c[j], be.scale(pk.sigma_cosets[j], beta)), gamma))
            dx = be.scale(pk.coset_x, beta * pow(DELTA, j, R_MOD) % R_MOD)
            right = be.mul(right, be.add_scalar(be.add(adv_c[j], dx), gamma))
        val = be.add(be.scale(val, y), be.mul(be.sub(left, right), pk.l_active_row))
    return val


def t_evaluations(k: int, ek: int):
    """1 / ((zeta omega_ext^i)^n - 1) for i < 2^(ek - k) (domain.rs:118-141)."""
    n = 1 << k
    w = omega_for(ek)
    return [pow((pow(ZETA * pow(w, i, R_MOD) % R_MOD, n, R_MOD) - 1) % R_MOD, -1, R_MOD)
            for i in range(1 << (ek - k))]


# Requests a proof yields to whoever drives it; the driver answers each with a list of
# results in order.  ("commit", [(srs, vector)]) -> affine points;
# ("ntt", [(op, vector, k, extended_k)]) -> vectors; ("eval", [(vector, x)]) -> ints.
def prove_steps(be, params: Params, pk: ProvingKey, advice_values, rng_seed: int,
                timings: dict | None = None):
    """create_proof (prover.rs:432-474) for the synthetic circuit, as a generator of
    commit / NTT / evaluation requests; returns the proof bytes.  advice_values are the
    five Lagrange columns from SyntheticCircuit.witness (rows >= n - 7 are replaced with
    blinding values)."""
    T0 = time.perf_counter()
    k, n = pk.k, pk.n
    ek = pk.dom.extended_k
    scale = pk.ext_n // n
    bf = BLINDING_FACTORS
    u = n - bf - 1
    tr = Transcript()
    rng_off = 0
    g_mono, g_lag = params.g, params.g_lagrange

    def draws(m):
        nonlocal rng_off
        v = be.from_seed(rng_seed, RNG_STREAM, m, rng_off)
        rng_off += m
        return v

    def mark(name):
        if timings is not None:
            be.sync()
            timings[name] = time.perf_counter() - T0

    # VK hash, instances (none), advice (prover.rs:101-108, 541-631)
    tr.common_fr(pk.transcript_repr)
    advice = [be.cat([be.slice(v, 0, u), draws(n - u)]) for v in advice_values]
    for c in (yield ("commit", [(g_lag, v) for v in advice])):
        tr.write_point(g1_compress(c))
    mark("advice_commit")
    theta = tr.squeeze()  # noqa: F841  (no lookups in this circuit)
    beta = tr.squeeze()
    gamma = tr.squeeze()
    perm_blind = [draws(bf) for _ in range(2)]

    # permutation grand products (permutation/prover.rs:81-255): each set's numerator and
    # denominator products (:124-164), one inversion of all denominators, their quotients
    w = be.powers(omega_for(k), n)
    sets = [list(range(0, CHUNK_LEN)), list(range(CHUNK_LEN, NUM_ADVICE))]
    specs = [[(advice[j], pk.sigma_values[j], pow(DELTA, j, R_MOD) * beta % R_MOD) for j in cols] for cols in sets]
    fused = getattr(be, "perm_products", None)
    num_den = fused(w, beta, gamma, specs) if fused else perm_products_composed(be, w, beta, gamma, specs)
    den_inv = be.batch_invert(be.cat([d for _, d in num_den]))
    zs = []
    last_uncorrected = []
    for s, (num, _) in enumerate(num_den):
        mv = be.mul(num, be.slice(den_inv, s * n, (s + 1) * n))
        # z_0 = 1, z_i = z_(i-1) * mv_(i-1): exclusive prefix products
        z = be.cat([be.const(1, 1), be.prefix_product(be.slice(mv, 0, n - 1))])
        z = be.cat([be.slice(z, 0, n - bf), perm_blind[s]])
        last_uncorrected.append(be.slice(z, u, u + 1))
        zs.append(z)
    # chain the sets: set s is multiplied by the product of the earlier sets' last values
    corr = None
    for s in range(1, len(zs)):
        corr = last_uncorrected[s - 1] if corr is None else be.mul(corr, last_uncorrected[s - 1])
        zs[s] = be.mul(zs[s], be.broadcast(corr, n))
    # z has long constant runs: committed as z.to_delta() against the suffix-sum bases
    # (permutation/prover.rs:236-245), the same point with the runs' zero deltas skipped
    for c in (yield ("commit", [(params.g_lagrange_delta, be.to_delta(z)) for z in zs])):
        tr.write_point(g1_compress(c))
    mark("permutation_commit")
    trash = tr.squeeze()  # noqa: F841  (no trashcans)
    y = tr.squeeze()
    polys = yield ("ntt", [("lagrange_to_coeff", v, k, ek) for v in advice + zs])
    advice_polys, z_polys = polys[:NUM_ADVICE], polys[NUM_ADVICE:]
    mark("advice_ifft")

    # quotient numerator on the extended coset (evaluation.rs:818-940), divided by X^n - 1
    cosets = yield ("ntt", [("coeff_to_extended", p, k, ek) for p in advice_polys + z_polys])
    adv_c, z_c = cosets[:NUM_ADVICE], cosets[NUM_ADVICE:]
    mark("coset_ntt")
    fused = getattr(be, "quotient_numerator", None)
    if fused is not None:
        h_ext = fused(pk, adv_c, z_c, y, beta, gamma)
    else:
        h_ext = be.divide_by_vanishing(numerator_composed(be, pk, adv_c, z_c, y, beta, gamma), k, ek)
    del adv_c, z_c, cosets
    mark("numerator")

    # h = nu / (X^n - 1), limbs (prover.rs:701-772)
    (h,) = yield ("ntt", [("extended_to_coeff", h_ext, k, ek)])
    del h_ext
    h = be.slice(h, 0, (n - 1) * QUOTIENT_DEGREE)
    limb_blind = draws(QUOTIENT_DEGREE - 1)
    limbs = []
    for i in range(QUOTIENT_DEGREE):
        body = be.slice(h, i * (n - 1), (i + 1) * (n - 1))
        tail = be.slice(limb_blind, i, i + 1) if i < QUOTIENT_DEGREE - 1 else be.zeros(1)
        limb = be.cat([body, tail])
        if i > 0:
            first = be.sub(be.slice(limb, 0, 1), be.slice(limb_blind, i - 1, i))
            limb = be.cat([first, be.slice(limb, 1, n)])
        limbs.append(limb)
    for c in (yield ("commit", [(g_mono, l) for l in limbs])):
        tr.write_point(g1_compress(c))
    mark("quotient_commit")
    x = tr.squeeze()

    # evaluations (prover.rs:798-869, permutation/prover.rs:266-340)
    x_next = _rotate(x, 1, k)
    x_last = _rotate(x, -(bf + 1), k)
    evals = yield ("eval", [(p, x) for p in advice_polys] + [(p, x) for p in pk.fixed_polys]
                   + [(p, x) for p in pk.sigma_polys]
                   + [(z_polys[0], x), (z_polys[0], x_next), (z_polys[0], x_last),
                      (z_polys[1], x), (z_polys[1], x_next)])
    for e in evals:
        tr.write_fr(e)
    mark("evaluations")

    # linearization (linearization/prover.rs:42-86): no simple selectors here, so every
    # identity is fully evaluated and L'(X) = -(x^n - 1) sum_j (x^(n-1))^j h_j(X)
    split = pow(x, n - 1, R_MOD)
    xn = split * x % R_MOD
    sp = (xn - 1) % R_MOD
    coeffs = []
    for _ in limbs:
        coeffs.append((-sp) % R_MOD)
        sp = sp * split % R_MOD
    lin = be.inner_product(limbs, coeffs)

    # queries in compute_queries order (prover.rs:871-944)
    queries = [(("advice", i), x, advice_polys[i]) for i in range(NUM_ADVICE)]
    queries += [(("perm", 0), x, z_polys[0]), (("perm", 0), x_next, z_polys[0]),
                (("perm", 1), x, z_polys[1]), (("perm", 1), x_next, z_polys[1]),
                (("perm", 0), x_last, z_polys[0])]
    queries += [(("fixed", i), x, pk.fixed_polys[i]) for i in range(NUM_FIXED)]
    queries += [(("sigma", i), x, pk.sigma_polys[i]) for i in range(NUM_ADVICE)]
    queries += [(("lin", 0), x, lin)]
    yield from multi_open(be, g_mono, queries, tr)
    mark("multi_open")
    return bytes(tr.proof)


def serve_request(be, req):
    """Answers one request of prove_steps on backend be."""
    kind, items = req
    if kind == "commit":
        return be.commit_jobs(items)
    if kind == "ntt":
        return be.ntt_many(items)
    if kind == "eval":
        return be.eval_many(items)
    raise ValueError(f"unknown request {kind!r}")


def create_proof(be, params: Params, pk: ProvingKey, advice_values, rng_seed: int,
                 timings: dict | None = None) -> bytes:
    """create_proof (prover.rs:432-474): one proof, every request answered at once."""
    gen = prove_steps(be, params, pk, advice_values, rng_seed, timings)
    resp = None
    while True:
        try:
            req = gen.send(resp)
        except StopIteration as stop:
            return stop.value
        resp = serve_request(be, req)


def create_proofs_native(be, params: Params, pk: ProvingKey, advices, rng_seeds, phase_ms: list | None = None):
    """Proofs of one size in lockstep by the library's native prover (mzk_plonk_prove):
    the same proofs create_proof makes, one batched MSM / NTT / evaluation call per
    phase for the whole group.  advices: one five-column witness per proof."""
    import ctypes
    mzk = be.mzk
    key = mzk.PlonkKey()
    key.k = pk.k
    key.g = params.g.h.value if hasattr(params.g.h, "value") else params.g.h
    key.g_lagrange = params.g_lagrange.h.value if hasattr(params.g_lagrange.h, "value") else params.g_lagrange.h
    p = be._p
    for j in range(NUM_FIXED):
        key.fixed_polys[j] = p(pk.fixed_polys[j])
        key.fixed_cosets[j] = p(pk.fixed_cosets[j])
    for j in range(NUM_ADVICE):
        key.sigma_values[j] = p(pk.sigma_values[j])
        key.sigma_polys[j] = p(pk.sigma_polys[j])
        key.sigma_cosets[j] = p(pk.sigma_cosets[j])
    key.l0, key.l_last, key.l_active = p(pk.l0), p(pk.l_last), p(pk.l_active_row)
    key.coset_x, key.omega_powers = p(pk.coset_x), p(pk.omega_pows)
    key.transcript_repr[:] = [int(w) for w in mont_limbs(pk.transcript_repr)]
    gd = params.g_lagrange_delta.h
    key.g_lagrange_delta = gd.value if hasattr(gd, "value") else gd
    advices = list(advices)
    m = len(advices)
    cols = []
    for adv in advices:
        if len(adv) != NUM_ADVICE:
            raise ValueError("each witness has five columns")
        cols += [p(v) for v in adv]
    arr = (ctypes.c_void_p * max(1, len(cols)))(*cols)
    seeds = np.array([int(s) & M64 for s in rng_seeds], dtype=np.uint64)
    if len(seeds) != m:
        raise ValueError("one rng seed per proof")
    out = np.zeros(max(1, m) * mzk.PLONK_PROOF_BYTES, dtype=np.uint8)
    ph = np.zeros(8, dtype=np.float64)
    be.sync()
    mzk._check(be._L.mzk_plonk_prove(be.ctx.h, ctypes.addressof(key), m, ctypes.cast(arr, ctypes.c_void_p),
                                     seeds.ctypes.data, out.ctypes.data,
                                     ph.ctypes.data if phase_ms is not None else None))
    if phase_ms is not None:
        phase_ms[:] = list(ph)
    return [bytes(out[i * mzk.PLONK_PROOF_BYTES:(i + 1) * mzk.PLONK_PROOF_BYTES]) for i in range(m)]


def perm_products_composed(be, w, beta: int, gamma: int, specs):
    """Per set: (prod_j (value_j + delta_beta_j w + gamma), prod_j (value_j + beta sigma_j + gamma))
    from the backend's vector operations (permutation/prover.rs:129-160); specs holds, per
    set, the (value, sigma, delta_beta) of its columns."""
    out = []
    for cols in specs:
        num = den = None
        for value, sigma, db in cols:
            t = be.add_scalar(be.add(be.scale(sigma, beta), value), gamma)
            den = t if den is None else be.mul(den, t)
            u = be.add_scalar(be.add(be.scale(w, db), value), gamma)
            num = u if num is None else be.mul(num, u)
        out.append((num, den))
    return out


def construct_intermediate_sets(queries):
    """kzg/utils.rs:40-133: (labels in first-appearance order with their set index,
    point sets with points in first-appearance order)."""
    point_index = {}
    com_labels, com_points = [], []
    for label, pt, _ in queries:
        pi = point_index.setdefault(pt, len(point_index))
        if label in com_labels:
            pos = com_labels.index(label)
            if pi in com_points[pos]:
                raise ValueError("duplicated query")
            com_points[pos].append(pi)
        else:
            com_labels.append(label)
            com_points.append([pi])
    set_index = {}
    com_sets = []
    for pts in com_points:
        key = tuple(sorted(pts))
        set_index.setdefault(key, len(set_index))
        com_sets.append(set_index[key])
    inv_point = {i: p for p, i in point_index.items()}
    point_sets = [None] * len(set_index)
    for key, si in set_index.items():
        point_sets[si] = [inv_point[i] for i in key]
    return com_labels, com_sets, point_sets


def multi_open(be, g_mono, queries, tr: Transcript):
    """KZGCommitmentScheme::multi_open (kzg/mod.rs:110-289), as a request generator."""
    x1 = tr.squeeze()
    x2 = tr.squeeze()
    label_poly = {}
    for label, _, poly in queries:
        label_poly[label] = poly
    labels, sets_of, point_sets = construct_intermediate_sets(queries)
    members = [[] for _ in point_sets]
    for label, si in zip(labels, sets_of):
        members[si].append(label_poly[label])
    q_polys = [be.inner_product(m, _powers(x1, len(m))) for m in members]
    order = sorted(range(len(point_sets)), key=lambda i: (len(point_sets[i]), i))
    q_polys = [q_polys[i] for i in order]
    point_sets = [point_sets[i] for i in order]
    f_parts = []
    for pts, q in zip(point_sets, q_polys):
        p = q
        for pt in pts:
            p = be.kate_div(p, pt)
        f_parts.append(p)
    f_poly = be.inner_product(f_parts, _powers(x2, len(f_parts)))
    (f_com,) = yield ("commit", [(g_mono, f_poly)])
    tr.write_point(g1_compress(f_com))
    x3 = tr.squeeze()
    for e in (yield ("eval", [(q, x3) for q in q_polys])):
        tr.write_fr(e)
    x4 = tr.squeeze()
    final = be.inner_product(q_polys + [f_poly], _powers(x4, len(q_polys) + 1))
    (v,) = yield ("eval", [(final, x3)])
    pi = be.kate_div(be.sub_scalar_at0(final, v), x3)
    (pi_com,) = yield ("commit", [(g_mono, pi)])
    tr.write_point(g1_compress(pi_com))


# ---------------------------------------------------------------- GPU backend
def _on_stream(cls):
    """Runs every public method of a GPU-side class with torch's current stream set to
    the library context's stream, so torch's copies and the HIP kernels are ordered."""
    for name, fn in list(vars(cls).items()):
        if name.startswith("_") or not callable(fn):
            continue

        def wrap(fn):
            def run(self, *a, **kw):
                with self._stream_ctx():
                    return fn(self, *a, **kw)
            run.__name__, run.__doc__ = fn.__name__, fn.__doc__
            return run
        setattr(cls, name, wrap(fn))
    return cls


@_on_stream
class GpuDomain:
    def __init__(self, be, j: int, k: int):
        self.be = be
        self.d = be.mzk.EvaluationDomain(be.ctx, j, k)
        self.k = self.d.k
        self.extended_k = self.d.extended_k

    def _stream_ctx(self):
        return self.be._stream_ctx()

    def lagrange_to_coeff(self, v):
        out = v.clone()
        self.d.lagrange_to_coeff(out)
        return out

    def coeff_to_extended(self, v):
        out = self.be._empty(1 << self.extended_k)
        self.d.coeff_to_extended(v, out)
        return out

    def extended_to_coeff(self, v):
        out = v.clone()
        self.d.extended_to_coeff(out)
        return out

    def divide_by_vanishing_poly(self, v):
        out = v.clone()
        self.d.divide_by_vanishing_poly(out)
        return out


@_on_stream
class GpuBackend:
    """Device vectors are torch int64 tensors of shape (len, 4) holding Montgomery
    limbs; every field operation is a HIP kernel of libmzk_gpu.so on the context's
    stream, and torch's own copies (concatenation, gather, roll) are queued on that same
    stream, so no operation needs a host wait."""

    def __init__(self, ctx=None, device: int = 0):
        import torch
        self.torch = torch
        from . import mzk
        self.mzk = mzk
        self.ctx = ctx if ctx is not None else self.mzk.Context(device)
        self.dev = torch.device("cuda", device)
        self.stream = torch.cuda.ExternalStream(self.ctx.stream, device=self.dev)
        self._L = self.mzk.load_library()

    # -- plumbing
    def _stream_ctx(self):
        return self.torch.cuda.stream(self.stream)

    def _empty(self, n):
        with self._stream_ctx():
            return self.torch.empty((n, 4), dtype=self.torch.int64, device=self.dev)

    def sync(self):
        self.ctx.synchronize()

    def mem_info(self):
        """(free, total) device bytes, for admission by HBM budget."""
        return self.ctx.mem_info()

    def zeros(self, n):
        with self._stream_ctx():
            return self.torch.zeros((n, 4), dtype=self.torch.int64, device=self.dev)

    def const(self, value: int, n: int):
        row = self.torch.from_numpy(mont_limbs(value).view(np.int64).copy())
        with self._stream_ctx():
            return row.to(self.dev, non_blocking=False).reshape(1, 4).repeat(n, 1)

    def broadcast(self, v1, n: int):
        with self._stream_ctx():
            return v1.reshape(1, 4).expand(n, 4).contiguous()

    def from_seed(self, seed: int, stream: int, n: int, offset: int = 0):
        out = self._empty(n)
        if n:
            self.ctx.fr_from_seed_dev(seed, stream, n, out, offset)
        return out

    def cat(self, vs):
        with self._stream_ctx():
            return self.torch.cat(list(vs), 0)

    def slice(self, v, lo, hi):
        return v[lo:hi]

    def gather(self, v, idx):
        with self._stream_ctx():
            return v.index_select(0, self.torch.as_tensor(np.asarray(idx, dtype=np.int64)).to(self.dev))

    def scatter(self, v, idx, src):
        with self._stream_ctx():
            out = v.clone()
            out[self.torch.as_tensor(np.asarray(idx, dtype=np.int64)).to(self.dev)] = src
            return out

    def select(self, mask, a, b):
        with self._stream_ctx():
            m = self.torch.as_tensor(np.asarray(mask, dtype=bool)).to(self.dev).reshape(-1, 1)
            return self.torch.where(m, a, b)

    def roll(self, v, rot: int):
        """w[i] = v[(i + rot) mod len] (a rotation of the evaluation domain)."""
        with self._stream_ctx():
            return self.torch.roll(v, -rot, 0)

    def to_host(self, v) -> np.ndarray:
        self.sync()
        return v.cpu().numpy().view(np.uint64).reshape(-1, 4)

    # -- arithmetic
    def _p(self, t):
        return t.data_ptr()

    def _binop(self, op, a, b):
        out = self._empty(a.shape[0])
        self.mzk._check(self._L.mzk_fr_vec_op(self.ctx.h, op, self._p(a), self._p(b), self._p(out), a.shape[0]))
        return out

    def add(self, a, b):
        return self._binop(0, a, b)

    def sub(self, a, b):
        return self._binop(1, a, b)

    def mul(self, a, b):
        return self._binop(2, a, b)

    def _scalar(self, op, a, s: int, out=None):
        out = self._empty(a.shape[0]) if out is None else out
        sl = mont_limbs(s)
        self.mzk._check(self._L.mzk_fr_vec_scalar(self.ctx.h, op, self._p(a), sl.ctypes.data, self._p(out),
                                                  a.shape[0]))
        return out

    def scale(self, a, s: int):
        return self._scalar(0, a, s)

    def add_scalar(self, a, s: int):
        return self._scalar(1, a, s)

    def neg(self, a):
        return self._scalar(0, a, R_MOD - 1)

    def axpy(self, acc, a, s: int):
        """acc + a * s (a may be shorter than acc: it updates acc's leading rows)."""
        out = acc.clone() if acc.shape[0] == a.shape[0] else acc.clone()
        self._scalar(2, a, s, out[:a.shape[0]])
        return out

    def inner_product(self, polys, scalars):
        """poly_inner_product (kzg/mod.rs:125-147): sum_i s_i p_i, zero-extended, one
        launch per 64 polynomials."""
        import ctypes
        polys, scalars = list(polys), list(scalars)[:len(polys)]
        m = max(p.shape[0] for p in polys)
        out = None
        for lo in range(0, len(polys), 64):
            part, sc = polys[lo:lo + 64], scalars[lo:lo + 64]
            q = len(part)
            ptrs = (ctypes.c_void_p * q)(*[self._p(p) if p.shape[0] else None for p in part])
            lens = (ctypes.c_size_t * q)(*[p.shape[0] for p in part])
            ss = np.concatenate([mont_limbs(x) for x in sc])
            res = self._empty(m)
            self.mzk._check(self._L.mzk_fr_lincomb_dev(self.ctx.h, ctypes.cast(ptrs, ctypes.c_void_p),
                                                       ctypes.cast(lens, ctypes.c_void_p), ss.ctypes.data, q,
                                                       self._p(res), m))
            out = res if out is None else self.add(out, res)
        return out

    def quotient_numerator(self, pk, adv_c, z_c, y: int, beta: int, gamma: int):
        """The numerator of the quotient on the extended coset, divided by X^n - 1, in one
        kernel (mzk_plonk_numerator_dev); equal to numerator_composed followed by
        divide_by_vanishing."""
        mzk = self.mzk
        A = mzk.PlonkNumerator()
        k, ek = pk.k, pk.dom.extended_k
        A.n_ext = pk.ext_n
        A.rot_next = pk.ext_n // pk.n
        A.rot_last = (BLINDING_FACTORS + 1) * (pk.ext_n // pk.n)
        A.ncols, A.nsets, A.chunk = NUM_ADVICE, 2, CHUNK_LEN
        te = t_evaluations(k, ek)
        A.t_count = len(te)
        A.a, A.b, A.c = self._p(adv_c[0]), self._p(adv_c[1]), self._p(adv_c[2])
        A.qm, A.ql, A.qr, A.qo, A.qc = (self._p(v) for v in pk.fixed_cosets)
        for j in range(NUM_ADVICE):
            A.col[j] = self._p(adv_c[j])
            A.sigma[j] = self._p(pk.sigma_cosets[j])
            A.delta_beta[j][:] = [int(w) for w in mont_limbs(beta * pow(DELTA, j, R_MOD) % R_MOD)]
        for s in range(2):
            A.z[s] = self._p(z_c[s])
        A.l0, A.l_last, A.l_active, A.coset_x = (self._p(pk.l0), self._p(pk.l_last), self._p(pk.l_active_row),
                                                 self._p(pk.coset_x))
        A.y[:] = [int(w) for w in mont_limbs(y)]
        A.beta[:] = [int(w) for w in mont_limbs(beta)]
        A.gamma[:] = [int(w) for w in mont_limbs(gamma)]
        for i, t in enumerate(te):
            A.t_evals[i][:] = [int(w) for w in mont_limbs(t)]
        out = self._empty(pk.ext_n)
        import ctypes
        mzk._check(self._L.mzk_plonk_numerator_dev(self.ctx.h, ctypes.addressof(A), self._p(out)))
        return out

    def sub_scalar_at0(self, a, v: int):
        """a(X) - v (subtracts from the constant coefficient)."""
        out = a.clone()
        self._scalar(1, a[:1], (-v) % R_MOD, out[:1])
        return out

    def perm_products(self, w, beta: int, gamma: int, specs):
        """perm_products_composed in one fused pass per set (mzk_plonk_perm_products_dev)."""
        import ctypes
        n = w.shape[0]
        sets, outs = [], []
        for cols in specs:
            e = self.mzk.PermSet()
            for j, (value, sigma, db) in enumerate(cols):
                e.values[j] = self._p(value)
                e.sigma[j] = self._p(sigma)
                e.delta_beta[j][:] = [int(x) for x in mont_limbs(db)]
            e.beta[:] = [int(x) for x in mont_limbs(beta)]
            e.gamma[:] = [int(x) for x in mont_limbs(gamma)]
            e.ncols = len(cols)
            num, den = self._empty(n), self._empty(n)
            e.num, e.den = self._p(num), self._p(den)
            sets.append(e)
            outs.append((num, den))
        arr = (self.mzk.PermSet * len(sets))(*sets)
        self.mzk._check(self._L.mzk_plonk_perm_products_dev(self.ctx.h, self._p(w), n, ctypes.cast(arr, ctypes.c_void_p),
                                                            len(sets)))
        return outs

    def batch_invert(self, a):
        out = a.clone()
        self.mzk._check(self._L.mzk_fr_batch_invert(self.ctx.h, self._p(out), out.shape[0]))
        return out

    def prefix_product(self, a):
        out = self._empty(a.shape[0])
        self.mzk._check(self._L.mzk_fr_prefix_product(self.ctx.h, self._p(a), self._p(out), a.shape[0]))
        return out

    def powers(self, x: int, n: int):
        out = self._empty(n)
        xl = mont_limbs(x)
        self.mzk._check(self._L.mzk_fr_powers(self.ctx.h, xl.ctypes.data, n, self._p(out)))
        return out

    def eval_many(self, items):
        """sum_i p_i x^i for every (p, x): one launch and one wait for all of them."""
        import ctypes
        items = list(items)
        if not items:
            return []
        m = len(items)
        ptrs = (ctypes.c_void_p * m)(*[self._p(p) if p.shape[0] else None for p, _ in items])
        ns = (ctypes.c_size_t * m)(*[p.shape[0] for p, _ in items])
        xs = np.concatenate([mont_limbs(x) for _, x in items])
        outs = np.zeros(4 * m, dtype=np.uint64)
        self.mzk._check(self._L.mzk_fr_eval_batch(self.ctx.h, ctypes.cast(ptrs, ctypes.c_void_p),
                                                  ctypes.cast(ns, ctypes.c_void_p), xs.ctypes.data, m,
                                                  outs.ctypes.data))
        return [from_mont_limbs(outs[4 * j:4 * j + 4]) for j in range(m)]

    def kate_div(self, a, b: int):
        q = self._empty(max(0, a.shape[0] - 1))
        bl = mont_limbs(b)
        self.mzk._check(self._L.mzk_fr_kate_div(self.ctx.h, self._p(a), a.shape[0], bl.ctypes.data,
                                                self._p(q) if q.shape[0] else None))
        return q

    def kate_div_many(self, items):
        """[(a, b)] -> the quotients of a(X) - a(b) by (X - b), one batched call."""
        import ctypes
        qs = [self._empty(max(0, a.shape[0] - 1)) for a, _ in items]
        jobs = (self.mzk.KateJob * max(1, len(items)))()
        for j, ((a, b), q) in enumerate(zip(items, qs)):
            jobs[j].d_a = self._p(a)
            jobs[j].n = a.shape[0]
            jobs[j].b[:] = [int(w) for w in mont_limbs(b)]
            jobs[j].d_q = self._p(q) if q.shape[0] else None
        self.mzk._check(self._L.mzk_fr_kate_div_batch(self.ctx.h, ctypes.addressof(jobs), len(items)))
        return qs

    # -- domains and commitments
    def domain(self, j: int, k: int):
        return GpuDomain(self, j, k)

    def srs_monomial(self, tau: int, n: int):
        return self.mzk.ResidentSrs.generate(self.ctx, mont_limbs(tau), n)

    def srs_from_scalars(self, s):
        import ctypes
        h = ctypes.c_void_p()
        self.mzk._check(self._L.mzk_srs_from_scalars_dev(self.ctx.h, self._p(s), s.shape[0], ctypes.byref(h)))
        return self.mzk.ResidentSrs(self.ctx, h)

    def srs_suffix_sum(self, srs):
        """suffix_sum (kzg/params.rs:47-57): the delta bases of srs."""
        return srs.suffix_sum()

    def to_delta(self, v, order: int = 1):
        """Polynomial::to_delta / to_double_delta (poly/mod.rs:290-335)."""
        out = self._empty(v.shape[0])
        self.mzk._check(self._L.mzk_fr_to_delta_dev(self.ctx.h, self._p(v), self._p(out), v.shape[0], order))
        return out

    def divide_by_vanishing(self, v, k: int, ek: int):
        """divide_by_vanishing_poly (domain.rs:361-380) on this backend's stream."""
        doms = self.__dict__.setdefault("_doms", {})
        if (k, ek) not in doms:
            doms[(k, ek)] = self.mzk.EvaluationDomain(self.ctx, 1 + (1 << (ek - k)), k)
        out = v.clone()
        doms[(k, ek)].divide_by_vanishing_poly(out)
        return out

    def ntt_many(self, items):
        """EvaluationDomain transforms of any sizes in one batched launch pair
        (lagrange_to_coeff, coeff_to_extended, extended_to_coeff; domain.rs:249-328)."""
        mzk = self.mzk
        jobs, outs = [], []
        for op, v, k, ek in items:
            n, N = 1 << k, 1 << ek
            if op == "lagrange_to_coeff":
                out = self._empty(n)
                jobs.append(mzk.NttJob.make(v, out, k, mzk.NTT_INVERSE, n))
            elif op == "coeff_to_extended":
                out = self._empty(N)
                jobs.append(mzk.NttJob.make(v, out, ek, mzk.NTT_COSET, n))
            elif op == "extended_to_coeff":
                out = self._empty(N)
                jobs.append(mzk.NttJob.make(v, out, ek, mzk.NTT_INVERSE | mzk.NTT_COSET, N))
            else:
                raise ValueError(f"unknown transform {op!r}")
            if v.shape[0] != (N if op == "extended_to_coeff" else n):
                raise ValueError(f"{op}: input length {v.shape[0]} does not match the domain")
            outs.append(out)
        if jobs:
            self.ctx.ntt_batch(jobs)
        return outs

    def commit_jobs(self, items):
        """KZG commits (kzg/mod.rs:82-108): [(srs, coefficients)] -> affine points; all
        commits against one SRS go through one batched MSM pipeline."""
        out = [None] * len(items)
        groups = {}
        for i, (srs, v) in enumerate(items):
            if v.shape[0] > srs.n:
                raise ValueError(f"polynomial of length {v.shape[0]} exceeds the {srs.n} SRS bases")
            groups.setdefault(id(srs), (srs, []))[1].append(i)
        for srs, idx in groups.values():
            jac = srs.commit_batch([(items[i][1], items[i][1].shape[0], 0) for i in idx], ctx=self.ctx)
            for r, i in enumerate(idx):
                out[i] = self.mzk.g1_to_affine(jac[r])
        return out

    def commit_many(self, params: Params, items):
        """[(coefficients, "monomial" | "lagrange")] against params' bases."""
        return self.commit_jobs([(params.g if b == "monomial" else params.g_lagrange, v) for v, b in items])
