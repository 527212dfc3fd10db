"""The prover backend on the oracle's C restatement (test infrastructure).

Same methods as the product's GpuBackend, with numpy (len, 4) uint64 Montgomery rows
as vectors and the oracle for every field operation, NTT and MSM.  Running the
product's create_proof over this backend gives the CPU prover whose proof bytes the
GPU prover must reproduce exactly.
"""
import numpy as np

from tests.helpers import pkg
from oracle import cpu_ref as C

plonk = pkg.plonk
R_MOD = plonk.R_MOD
mont_limbs = plonk.mont_limbs
from_mont_limbs = plonk.from_mont_limbs


class CpuSrs:
    def __init__(self, bases: np.ndarray):
        self.bases = bases
        self.n = bases.shape[0]


class CpuDomain:
    def __init__(self, j: int, k: int):
        from oracle import bls12_381 as O
        self.k = k
        self.extended_k = O.extended_k_for(j, k)

    def lagrange_to_coeff(self, v):
        return C.intt(v, self.k)

    def coeff_to_extended(self, v):
        return C.coeff_to_extended(v, self.k, self.extended_k)

    def extended_to_coeff(self, v):
        return C.extended_to_coeff(v, self.extended_k)

    def divide_by_vanishing_poly(self, v):
        return C.divide_by_vanishing_poly(v, self.k, self.extended_k)


class CpuBackend:
    def __init__(self, threads: int = 8):
        C.set_threads(threads)

    def sync(self):
        pass

    def zeros(self, n):
        return np.zeros((n, 4), dtype=np.uint64)

    def const(self, value: int, n: int):
        return np.tile(mont_limbs(value), (n, 1))

    def broadcast(self, v1, n: int):
        return np.tile(np.asarray(v1).reshape(1, 4), (n, 1))

    def from_seed(self, seed: int, stream: int, n: int, offset: int = 0):
        if n == 0:
            return self.zeros(0)
        return C.fr_from_seed(seed, stream, n, offset)

    def cat(self, vs):
        return np.ascontiguousarray(np.concatenate(list(vs), 0))

    def slice(self, v, lo, hi):
        return np.ascontiguousarray(v[lo:hi])

    def gather(self, v, idx):
        return np.ascontiguousarray(v[np.asarray(idx, dtype=np.int64)])

    def scatter(self, v, idx, src):
        out = v.copy()
        out[np.asarray(idx, dtype=np.int64)] = src
        return out

    def select(self, mask, a, b):
        return np.ascontiguousarray(np.where(np.asarray(mask, dtype=bool).reshape(-1, 1), a, b))

    def roll(self, v, rot: int):
        return np.ascontiguousarray(np.roll(v, -rot, 0))

    def to_host(self, v):
        return v

    def add(self, a, b):
        return C.fr_vec_op(0, a, b)

    def sub(self, a, b):
        return C.fr_vec_op(1, a, b)

    def mul(self, a, b):
        return C.fr_vec_op(2, a, b)

    def scale(self, a, s: int):
        return C.fr_vec_scalar(0, a, mont_limbs(s))

    def add_scalar(self, a, s: int):
        return C.fr_vec_scalar(1, a, mont_limbs(s))

    def neg(self, a):
        return C.fr_vec_scalar(0, a, mont_limbs(R_MOD - 1))

    def axpy(self, acc, a, s: int):
        out = acc.copy()
        head = np.ascontiguousarray(out[:a.shape[0]])
        C.fr_vec_scalar(2, a, mont_limbs(s), head)
        out[:a.shape[0]] = head
        return out

    def inner_product(self, polys, scalars):
        m = max(p.shape[0] for p in polys)
        acc = self.zeros(m)
        for p, s in zip(polys, scalars):
            acc = self.axpy(acc, p, s)
        return acc

    def sub_scalar_at0(self, a, v: int):
        out = a.copy()
        out[:1] = C.fr_vec_scalar(1, np.ascontiguousarray(a[:1]), mont_limbs((-v) % R_MOD))
        return out

    def batch_invert(self, a):
        return C.fr_batch_invert(a)

    def prefix_product(self, a):
        return C.fr_prefix_product(a)

    def powers(self, x: int, n: int):
        return C.fr_powers(mont_limbs(x), n)

    def eval_many(self, items):
        return [from_mont_limbs(C.poly_eval(np.ascontiguousarray(p), mont_limbs(x))) for p, x in items]

    def kate_div(self, a, b: int):
        return C.fr_kate_div(a, mont_limbs(b))

    def domain(self, j: int, k: int):
        return CpuDomain(j, k)

    def srs_monomial(self, tau: int, n: int):
        return CpuSrs(C.srs_bases(mont_limbs(tau), n))

    def srs_from_scalars(self, s):
        return CpuSrs(C.g1_mul_generator_many(s))

    def srs_suffix_sum(self, srs):
        return CpuSrs(C.g1_suffix_sum(srs.bases))

    def to_delta(self, v, order: int = 1):
        return C.fr_to_delta(v, order)

    def divide_by_vanishing(self, v, k, ek):
        return C.divide_by_vanishing_poly(v, k, ek)

    def ntt_many(self, items):
        outs = []
        for op, v, k, ek in items:
            if op == "lagrange_to_coeff":
                outs.append(C.intt(v, k))
            elif op == "coeff_to_extended":
                outs.append(C.coeff_to_extended(v, k, ek))
            elif op == "extended_to_coeff":
                outs.append(C.extended_to_coeff(v, ek))
            else:
                raise ValueError(op)
        return outs

    def commit_jobs(self, items):
        out = []
        for srs, v in items:
            n = v.shape[0]
            if n > srs.n:
                raise ValueError("polynomial longer than the SRS")
            jac = C.msm_specific(np.ascontiguousarray(v), np.ascontiguousarray(srs.bases[:n]))
            out.append(C.g1_to_affine(jac))
        return out

    def commit_many(self, params, items):
        return self.commit_jobs([(params.g if b == "monomial" else params.g_lagrange, v) for v, b in items])
