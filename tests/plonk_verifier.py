"""Trapdoor verifier for the synthetic circuit's proofs (test infrastructure).

Restates the verifier side of midnight-proofs for this circuit: the transcript
reads of plonk/verifier.rs, the identity evaluation of partially_evaluate_identities
(plonk/mod.rs:502-620) with the permutation expressions of permutation.rs:193-298,
the linearization commitment (linearization/verifier.rs), and
KZGCommitmentScheme::multi_prepare (kzg/mod.rs:291-489).  The final pairing check
e(pi, [tau]G2) = e(R, G2) of DualMSM::check (kzg/msm.rs:294-310) is replaced by the
equivalent G1 equation [tau] pi = R, which needs the SRS trapdoor tau (known here
because the parameters come from unsafe_setup).  Group arithmetic is the oracle's
textbook affine/Jacobian code; the transcript is Python's hashlib BLAKE2b.
"""
import hashlib

from oracle import bls12_381 as O

R = O.R_MOD
P = O.P_MOD


def omega(k):
    return pow(O.FR_ROOT_OF_UNITY, 1 << (32 - k), R)


def decompress(b: bytes):
    if len(b) != 48 or not (b[0] & 0x80):
        raise ValueError("not a compressed point")
    if b[0] & 0x40:
        if any(b[1:]) or (b[0] & 0x3F):
            raise ValueError("bad infinity encoding")
        return None
    x = int.from_bytes(bytes([b[0] & 0x1F]) + b[1:], "big")
    if x >= P:
        raise ValueError("x out of range")
    y2 = (x * x * x + 4) % P
    y = pow(y2, (P + 1) // 4, P)
    if y * y % P != y2:
        raise ValueError("x not on the curve")
    if bool(b[0] & 0x20) != (y > (P - 1) // 2):
        y = P - y
    return (x, y)


def affine_from_limbs(limbs12):
    vals = [int(w) for w in limbs12]
    if not any(vals):
        return None
    xm = sum(vals[i] << (64 * i) for i in range(6))
    ym = sum(vals[6 + i] << (64 * i) for i in range(6))
    return (O.fp_from_mont(xm), O.fp_from_mont(ym))


class Reader:
    def __init__(self, proof: bytes):
        self.buf = proof
        self.pos = 0
        self.h = hashlib.blake2b(digest_size=64, key=b"Domain separator for transcript")

    def common(self, data: bytes):
        self.h.update(bytes([1]) + data)

    def squeeze(self):
        self.h.update(bytes([0]))
        return int.from_bytes(self.h.digest(), "little") % R

    def point(self):
        b = self.buf[self.pos:self.pos + 48]
        self.pos += 48
        self.common(b)
        return decompress(b)

    def scalar(self):
        b = self.buf[self.pos:self.pos + 32]
        self.pos += 32
        v = int.from_bytes(b, "little")
        if len(b) != 32 or v >= R:
            raise ValueError("bad scalar")
        self.common(b)
        return v


def msm(terms):
    return O.g1_sum([O.g1_mul(pt, s % R) for s, pt in terms if pt is not None and s % R])


def verify(proof: bytes, k: int, transcript_repr: int, fixed_coms, perm_coms, tau: int) -> bool:
    n = 1 << k
    bf = 6
    w = omega(k)
    winv = pow(w, -1, R)
    DELTA = O.FR_DELTA
    tr = Reader(proof)
    tr.common(transcript_repr.to_bytes(32, "little"))
    adv = [tr.point() for _ in range(5)]
    tr.squeeze()                      # theta
    beta = tr.squeeze()
    gamma = tr.squeeze()
    zc = [tr.point() for _ in range(2)]
    tr.squeeze()                      # trash challenge
    y = tr.squeeze()
    hc = [tr.point() for _ in range(4)]
    x = tr.squeeze()
    adv_e = [tr.scalar() for _ in range(5)]
    fix_e = [tr.scalar() for _ in range(5)]
    sig_e = [tr.scalar() for _ in range(5)]
    z0, z0n, z0l = tr.scalar(), tr.scalar(), tr.scalar()
    z1, z1n = tr.scalar(), tr.scalar()

    split = pow(x, n - 1, R)
    xn = split * x % R
    # l_i(x) for i in -(bf+1)..=0 (domain.rs l_i_range)
    def l_i(i):
        wi = pow(w, i, R) if i >= 0 else pow(winv, -i, R)
        return (xn - 1) * pow(n, -1, R) % R * wi % R * pow((x - wi) % R, -1, R) % R
    l_last = l_i(-(bf + 1))
    l_blind = sum(l_i(-i) for i in range(1, bf + 1)) % R
    l_0 = l_i(0)
    a, b, c = adv_e[0], adv_e[1], adv_e[2]
    qm, ql, qr, qo, qc = fix_e
    exprs = [(a * b % R * qm + a * ql + b * qr - c * qo + qc) % R]
    exprs.append(l_0 * (1 - z0) % R)
    exprs.append((z1 * z1 - z1) * l_last % R)
    exprs.append((z1 - z0l) * l_0 % R)
    sets = [([0, 1, 2], z0, z0n), ([3, 4], z1, z1n)]
    for ci, (cols, zv, zn) in enumerate(sets):
        left = zn
        for j in cols:
            left = left * (adv_e[j] + beta * sig_e[j] + gamma) % R
        right = zv
        cur = beta * x % R * pow(DELTA, ci * 3, R) % R
        for j in cols:
            right = right * (adv_e[j] + cur + gamma) % R
            cur = cur * DELTA % R
        exprs.append((left - right) * (1 - (l_last + l_blind)) % R)
    const = 0
    ypow = 1
    for e in reversed(exprs):
        const = (const + ypow * e) % R
        ypow = ypow * y % R
    # the linearization commitment: -(x^n - 1) sum_j split^j H_j, evaluating to -C
    sp = (xn - 1) % R
    lin_terms = []
    for H in hc:
        lin_terms.append(((-sp) % R, H))
        sp = sp * split % R
    x_next = x * w % R
    x_last = x * pow(winv, bf + 1, R) % R
    # queries: (label, point, commitment terms [(scalar, point)], eval)
    Q = [(("advice", i), x, [(1, adv[i])], adv_e[i]) for i in range(5)]
    Q += [(("perm", 0), x, [(1, zc[0])], z0), (("perm", 0), x_next, [(1, zc[0])], z0n),
          (("perm", 1), x, [(1, zc[1])], z1), (("perm", 1), x_next, [(1, zc[1])], z1n),
          (("perm", 0), x_last, [(1, zc[0])], z0l)]
    Q += [(("fixed", i), x, [(1, fixed_coms[i])], fix_e[i]) for i in range(5)]
    Q += [(("sigma", i), x, [(1, perm_coms[i])], sig_e[i]) for i in range(5)]
    Q += [(("lin", 0), x, lin_terms, (-const) % R)]

    # multi_prepare
    x1 = tr.squeeze()
    x2 = tr.squeeze()
    pidx = {}
    labels, lpts, lcom, levals = [], [], {}, {}
    for lab, pt, com, ev in Q:
        pi = pidx.setdefault(pt, len(pidx))
        if lab not in labels:
            labels.append(lab)
            lpts.append([])
        lpts[labels.index(lab)].append(pi)
        lcom[lab] = com
        levals.setdefault(lab, {})[pi] = ev
    set_id = {}
    lab_set = []
    for pts in lpts:
        key = tuple(sorted(pts))
        set_id.setdefault(key, len(set_id))
        lab_set.append(set_id[key])
    inv = {i: p for p, i in pidx.items()}
    point_sets = [None] * len(set_id)
    for key, si in set_id.items():
        point_sets[si] = [inv[i] for i in key]
    set_keys = [None] * len(set_id)
    for key, si in set_id.items():
        set_keys[si] = key
    q_terms = [[] for _ in point_sets]
    q_evals = [[0] * len(ps) for ps in point_sets]
    counts = [0] * len(point_sets)
    for lab, si in zip(labels, lab_set):
        s = pow(x1, counts[si], R)
        counts[si] += 1
        q_terms[si] += [(s * t % R, pt) for t, pt in lcom[lab]]
        for m, pi in enumerate(set_keys[si]):
            q_evals[si][m] = (q_evals[si][m] + s * levals[lab][pi]) % R
    order = sorted(range(len(point_sets)), key=lambda i: (len(point_sets[i]), i))
    q_terms = [q_terms[i] for i in order]
    q_evals = [q_evals[i] for i in order]
    point_sets = [point_sets[i] for i in order]
    f_com = tr.point()
    x3 = tr.squeeze()
    q_x3 = [tr.scalar() for _ in point_sets]
    f_eval = 0
    for pts, evs, pe in reversed(list(zip(point_sets, q_evals, q_x3))):
        # r(x3) of the interpolating polynomial through (pts, evs)
        r_eval = 0
        for j, pj in enumerate(pts):
            num, den = 1, 1
            for m, pm in enumerate(pts):
                if m != j:
                    num = num * (x3 - pm) % R
                    den = den * (pj - pm) % R
            r_eval = (r_eval + evs[j] * num * pow(den, -1, R)) % R
        den = 1
        for p in pts:
            den = den * (x3 - p) % R
        f_eval = (f_eval * x2 + (pe - r_eval) * pow(den, -1, R)) % R
    x4 = tr.squeeze()
    final_terms = []
    pw = 1
    for terms in q_terms:
        final_terms += [(pw * s % R, pt) for s, pt in terms]
        pw = pw * x4 % R
    final_terms.append((pw, f_com))
    v = 0
    pw = 1
    for e in q_x3 + [f_eval]:
        v = (v + pw * e) % R
        pw = pw * x4 % R
    pi = tr.point()
    if tr.pos != len(proof):
        return False
    rhs = msm(final_terms + [((-v) % R, O.G1_GEN), (x3, pi)])
    lhs = O.g1_mul(pi, tau % R) if pi is not None else None
    return lhs == rhs
