"""ctypes view of the oracle's C restatement (oracle/libcpu_ref.so).

TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.  Used by
tests/, the smoke check and bench.py's cpu_baseline leg; never by the product.
Arrays are numpy uint64 in the blst limb layouts: fr = 4 u64 Montgomery,
G1 affine = 12 u64 (x || y), G1 Jacobian = 18 u64 (X || Y || Z).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpu_ref.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "cpu_ref.c")):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        U = ctypes.c_uint
        U64 = ctypes.c_uint64
        sig = {
            "cpu_ref_set_threads": ([ctypes.c_int], None),
            "cpu_ref_get_threads": ([], ctypes.c_int),
            "cpu_ref_booth_index": ([ctypes.c_int, ctypes.c_int, P], ctypes.c_int32),
            "cpu_ref_window_size": ([S], ctypes.c_int),
            "cpu_ref_msm_best": ([P, P, S, P], None),
            "cpu_ref_msm_parallel": ([P, P, S, P], None),
            "cpu_ref_msm_serial": ([P, P, S, P], None),
            "cpu_ref_msm_specific": ([P, P, S, P], None),
            "cpu_ref_g1_to_affine": ([P, P], None),
            "cpu_ref_compute_twiddles": ([P, U, P], None),
            "cpu_ref_best_fft_with_twiddles": ([P, P, U], None),
            "cpu_ref_fft_coeff_to_extended": ([P, P, U, S], None),
            "cpu_ref_root_of_unity": ([U, P], None),
            "cpu_ref_fr_inv": ([P, P], None),
            "cpu_ref_fr_mul": ([P, P, P], None),
            "cpu_ref_ntt": ([P, U], None),
            "cpu_ref_intt": ([P, U], None),
            "cpu_ref_coeff_to_extended": ([P, U, U, P], None),
            "cpu_ref_extended_to_coeff": ([P, U], None),
            "cpu_ref_divide_by_vanishing_poly": ([P, U, U], None),
            "cpu_ref_fr_from_seed": ([U64, U64, U64, S, P], None),
            "cpu_ref_g1_generator": ([P], None),
            "cpu_ref_srs_bases": ([P, S, P], None),
            "cpu_ref_poly_eval": ([P, S, P, P], None),
            "cpu_ref_g1_mul_generator": ([P, P], None),
            "cpu_ref_fr_vec_op": ([ctypes.c_int, P, P, P, S], None),
            "cpu_ref_fr_vec_scalar": ([ctypes.c_int, P, P, P, S], None),
            "cpu_ref_fr_batch_invert": ([P, S], None),
            "cpu_ref_fr_prefix_product": ([P, P, S], None),
            "cpu_ref_fr_kate_div": ([P, S, P, P], None),
            "cpu_ref_fr_powers": ([P, S, P], None),
            "cpu_ref_g1_mul_generator_many": ([P, S, P], None),
            "cpu_ref_g1_suffix_sum": ([P, S, P], None),
            "cpu_ref_fr_to_delta": ([P, S, ctypes.c_int, P], None),
            "cpu_ref_logup_multiplicities": ([P, P, S, P, S, S, P, P], ctypes.c_longlong),
            "cpu_ref_logup_logderivative": ([S, S, P, P, P, P, S, P, S, P, P, P, P], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(t: int):
    lib().cpu_ref_set_threads(int(t))


def fr_from_seed(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().cpu_ref_fr_from_seed(seed, stream, offset, n, _p(out))
    return out


def srs_bases(tau_mont: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros((n, 12), dtype=np.uint64)
    lib().cpu_ref_srs_bases(_p(np.ascontiguousarray(tau_mont, dtype=np.uint64)), n, _p(out))
    return out


def msm_best(scalars: np.ndarray, bases: np.ndarray) -> np.ndarray:
    out = np.zeros(18, dtype=np.uint64)
    lib().cpu_ref_msm_best(_p(scalars), _p(bases), scalars.shape[0], _p(out))
    return out


def msm_parallel(scalars: np.ndarray, bases: np.ndarray) -> np.ndarray:
    out = np.zeros(18, dtype=np.uint64)
    lib().cpu_ref_msm_parallel(_p(scalars), _p(bases), scalars.shape[0], _p(out))
    return out


def msm_serial(scalars: np.ndarray, bases: np.ndarray, acc: np.ndarray) -> np.ndarray:
    """acc + the MSM (msm.rs:338-423), Jacobian."""
    out = np.ascontiguousarray(acc, dtype=np.uint64).reshape(18).copy()
    lib().cpu_ref_msm_serial(_p(scalars), _p(bases), scalars.shape[0], _p(out))
    return out


def msm_specific(scalars: np.ndarray, bases: np.ndarray) -> np.ndarray:
    out = np.zeros(18, dtype=np.uint64)
    lib().cpu_ref_msm_specific(_p(scalars), _p(bases), scalars.shape[0], _p(out))
    return out


def g1_to_affine(jac: np.ndarray) -> np.ndarray:
    out = np.zeros(12, dtype=np.uint64)
    lib().cpu_ref_g1_to_affine(_p(np.ascontiguousarray(jac, dtype=np.uint64)), _p(out))
    return out


def g1_mul_generator(s_mont: np.ndarray) -> np.ndarray:
    out = np.zeros(18, dtype=np.uint64)
    lib().cpu_ref_g1_mul_generator(_p(np.ascontiguousarray(s_mont, dtype=np.uint64)), _p(out))
    return g1_to_affine(out)


def poly_eval(coeffs: np.ndarray, x_mont: np.ndarray) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    lib().cpu_ref_poly_eval(_p(coeffs), coeffs.shape[0], _p(np.ascontiguousarray(x_mont, dtype=np.uint64)), _p(out))
    return out


def root_of_unity(log_n: int) -> np.ndarray:
    """ROOT^(2^(32 - log_n)), Montgomery."""
    out = np.zeros(4, dtype=np.uint64)
    lib().cpu_ref_root_of_unity(log_n, _p(out))
    return out


def fr_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    lib().cpu_ref_fr_mul(_p(_c(a)), _p(_c(b)), _p(out))
    return out


def fr_inv(a: np.ndarray) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    lib().cpu_ref_fr_inv(_p(_c(a)), _p(out))
    return out


def compute_twiddles(omega: np.ndarray, log_n: int) -> np.ndarray:
    """[1, w, ..., w^(n/2 - 1)] (fft.rs:31-40)."""
    out = np.zeros((1 << (log_n - 1), 4), dtype=np.uint64)
    lib().cpu_ref_compute_twiddles(_p(_c(omega)), log_n, _p(out))
    return out


def best_fft_with_twiddles(a: np.ndarray, twiddles: np.ndarray, log_n: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_best_fft_with_twiddles(_p(b), _p(_c(twiddles)), log_n)
    return b


def fft_coeff_to_extended(a: np.ndarray, twiddles: np.ndarray, log_n: int, n_real: int) -> np.ndarray:
    """The pruned DIF of fft.rs:110-186 (a's entries from n_real on must be zero)."""
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_fft_coeff_to_extended(_p(b), _p(_c(twiddles)), log_n, n_real)
    return b


def ntt(a: np.ndarray, log_n: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_ntt(_p(b), log_n)
    return b


def intt(a: np.ndarray, log_n: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_intt(_p(b), log_n)
    return b


def coeff_to_extended(a: np.ndarray, k: int, ext_k: int) -> np.ndarray:
    out = np.zeros((1 << ext_k, 4), dtype=np.uint64)
    lib().cpu_ref_coeff_to_extended(_p(np.ascontiguousarray(a, dtype=np.uint64)), k, ext_k, _p(out))
    return out


def extended_to_coeff(a: np.ndarray, ext_k: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_extended_to_coeff(_p(b), ext_k)
    return b


def divide_by_vanishing_poly(a: np.ndarray, k: int, ext_k: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().cpu_ref_divide_by_vanishing_poly(_p(b), k, ext_k)
    return b


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def fr_vec_op(op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """op 0: a + b, 1: a - b, 2: a * b (element-wise, Montgomery rows)."""
    a, b = _c(a), _c(b)
    out = np.empty_like(a)
    lib().cpu_ref_fr_vec_op(op, _p(a), _p(b), _p(out), a.shape[0])
    return out


def fr_vec_scalar(op: int, a: np.ndarray, s_mont: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
    """op 0: a * s, 1: a + s, 2: out + a * s (out updated in place)."""
    a = _c(a)
    if out is None:
        out = np.empty_like(a)
    lib().cpu_ref_fr_vec_scalar(op, _p(a), _p(_c(s_mont)), _p(out), a.shape[0])
    return out


def fr_batch_invert(a: np.ndarray) -> np.ndarray:
    b = _c(a).copy()
    lib().cpu_ref_fr_batch_invert(_p(b), b.shape[0])
    return b


def fr_prefix_product(a: np.ndarray) -> np.ndarray:
    a = _c(a)
    out = np.empty_like(a)
    lib().cpu_ref_fr_prefix_product(_p(a), _p(out), a.shape[0])
    return out


def fr_kate_div(a: np.ndarray, b_mont: np.ndarray) -> np.ndarray:
    a = _c(a)
    q = np.zeros((max(0, a.shape[0] - 1), 4), dtype=np.uint64)
    lib().cpu_ref_fr_kate_div(_p(a), a.shape[0], _p(_c(b_mont)), _p(q))
    return q


def fr_powers(x_mont: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().cpu_ref_fr_powers(_p(_c(x_mont)), n, _p(out))
    return out


def g1_suffix_sum(bases: np.ndarray) -> np.ndarray:
    """suffix_sum (params.rs:47-57) of affine bases (n x 12)."""
    b = _c(bases)
    out = np.zeros_like(b)
    lib().cpu_ref_g1_suffix_sum(_p(b), b.shape[0], _p(out))
    return out


def fr_to_delta(a: np.ndarray, order: int = 1) -> np.ndarray:
    """Polynomial::to_delta (order 1) / to_double_delta (order 2), poly/mod.rs:290-335."""
    a = _c(a)
    out = np.empty_like(a)
    lib().cpu_ref_fr_to_delta(_p(a), a.shape[0], order, _p(out))
    return out


def _ptrs(arrays):
    return (ctypes.c_void_p * max(1, len(arrays)))(*[a.ctypes.data for a in arrays])


def logup_multiplicities(selector, inputs, table, usable: int, blinding) -> np.ndarray:
    """compute_multiplicities (logup/prover.rs:411-477).  Raises ValueError naming the
    first selected input value (column, row) missing from the table."""
    sel, tab = _c(selector), _c(table)
    ins = [_c(x) for x in inputs]
    n = tab.shape[0]
    bl = _c(blinding) if len(blinding) else np.zeros((1, 4), dtype=np.uint64)
    out = np.zeros((n, 4), dtype=np.uint64)
    bad = lib().cpu_ref_logup_multiplicities(_p(sel), _ptrs(ins), len(ins), _p(tab), n, usable, _p(bl), _p(out))
    if bad:
        c, i = divmod(bad - 1, n)
        raise ValueError(f"input {c} row {i}: value not found in the lookup table")
    return out


def logup_logderivative(selector, m, table, inputs, chunk_sizes, beta_mont, blinding, bf: int):
    """compute_logderivative (logup/prover.rs:189-310): (helper columns, aggregator)."""
    sel, mm, tab = _c(selector), _c(m), _c(table)
    n = tab.shape[0]
    ins = [_c(x) for x in inputs]
    cs = np.asarray(chunk_sizes, dtype=np.uint32)
    helpers = [np.zeros((n, 4), dtype=np.uint64) for _ in chunk_sizes]
    z = np.zeros((n, 4), dtype=np.uint64)
    bl = _c(blinding) if bf else np.zeros((1, 4), dtype=np.uint64)
    lib().cpu_ref_logup_logderivative(n, bf, _p(sel), _p(mm), _p(tab), _ptrs(ins), len(ins),
                                      cs.ctypes.data_as(ctypes.c_void_p), len(cs), _p(_c(beta_mont)), _p(bl),
                                      _ptrs(helpers), _p(z))
    return helpers, z


def g1_mul_generator_many(s_mont: np.ndarray) -> np.ndarray:
    s = _c(s_mont)
    out = np.zeros((s.shape[0], 12), dtype=np.uint64)
    lib().cpu_ref_g1_mul_generator_many(_p(s), s.shape[0], _p(out))
    return out


# ---- conversions between numpy limb arrays and Python ints (canonical)
def fr_rows_to_ints(a: np.ndarray):
    from . import bls12_381 as O
    out = []
    for row in np.asarray(a, dtype=np.uint64).reshape(-1, 4):
        out.append(O.fr_from_mont(O.from_limbs(int(v) for v in row)))
    return out


def ints_to_fr_rows(vals) -> np.ndarray:
    from . import bls12_381 as O
    out = np.zeros((len(vals), 4), dtype=np.uint64)
    for i, v in enumerate(vals):
        out[i] = O.to_limbs(O.fr_to_mont(v), 4)
    return out
