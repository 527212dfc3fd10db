"""Python host mirror of the reference's MSM / NTT interface over the C ABI.

Loads the in-tree gfx950 library libmzk_gpu.so (built by build.py) through ctypes
and exposes the reference's names:

  msm_best(coeffs, bases)            curves/src/msm.rs:460
  msm_parallel(coeffs, bases)        curves/src/msm.rs:430
  msm_serial(coeffs, bases, acc)     curves/src/msm.rs:338
  msm_specific(coeffs, bases)        proofs/src/poly/kzg/msm.rs:174
  compute_twiddles(omega, log_n)     curves/src/fft.rs:31
  best_fft(a, omega, log_n)          curves/src/fft.rs:49
  best_fft_with_twiddles(a, tw, log_n)          curves/src/fft.rs:56
  fft_coeff_to_extended(a, tw, log_n, n_real)   curves/src/fft.rs:110
  EvaluationDomain(j, k)             proofs/src/poly/domain.rs:54
     .lagrange_to_coeff / .coeff_to_lagrange / .coeff_to_extended /
     .extended_to_coeff / .divide_by_vanishing_poly
  ParamsKZG-style resident bases     proofs/src/poly/kzg/params.rs:22-37,154

Host arrays are numpy uint64 in the blst limb layouts (fr = 4 limbs, G1 affine =
12, G1 Jacobian = 18).  Device arrays are torch tensors (torch provides device
memory and nothing else).  Errors raise MzkError, mirroring the reference's
panics on bad lengths (msm.rs:461, fft.rs:62).  There is no CPU fallback: if the
library or a GPU is missing, calls fail loudly.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmzk_gpu.so")


class MzkError(RuntimeError):
    pass


class MsmJob(ctypes.Structure):
    """mzk_msm_job: one commitment of a batch."""
    _fields_ = [("d_coeffs", ctypes.c_void_p), ("n", ctypes.c_size_t), ("offset", ctypes.c_size_t)]


NTT_INVERSE = 1
NTT_COSET = 2
NTT_NOALLOC_MAX_LOG = 20


class NttJob(ctypes.Structure):
    """mzk_ntt_job: one transform of size 2^log_n (flags NTT_INVERSE / NTT_COSET)."""
    _fields_ = [("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("n_real", ctypes.c_uint64),
                ("log_n", ctypes.c_uint32), ("flags", ctypes.c_uint32)]

    @classmethod
    def make(cls, d_in, d_out, log_n: int, flags: int = 0, n_real: int | None = None) -> "NttJob":
        j = cls()
        j.d_in = _t_ptr(d_in)
        j.d_out = _t_ptr(d_out)
        j.log_n = log_n
        j.flags = flags
        j.n_real = (1 << log_n) if n_real is None else n_real
        return j


class PlonkNumerator(ctypes.Structure):
    """mzk_plonk_numerator: the fused quotient-numerator arguments."""
    _fields_ = [("n_ext", ctypes.c_uint64), ("rot_next", ctypes.c_uint32), ("rot_last", ctypes.c_uint32),
                ("ncols", ctypes.c_uint32), ("nsets", ctypes.c_uint32), ("chunk", ctypes.c_uint32),
                ("t_count", ctypes.c_uint32),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p), ("qm", ctypes.c_void_p),
                ("ql", ctypes.c_void_p), ("qr", ctypes.c_void_p), ("qo", ctypes.c_void_p), ("qc", ctypes.c_void_p),
                ("col", ctypes.c_void_p * 8), ("sigma", ctypes.c_void_p * 8), ("z", ctypes.c_void_p * 4),
                ("l0", ctypes.c_void_p), ("l_last", ctypes.c_void_p), ("l_active", ctypes.c_void_p),
                ("coset_x", ctypes.c_void_p),
                ("y", ctypes.c_uint64 * 4), ("beta", ctypes.c_uint64 * 4), ("gamma", ctypes.c_uint64 * 4),
                ("delta_beta", (ctypes.c_uint64 * 4) * 8), ("t_evals", (ctypes.c_uint64 * 4) * 16)]


PERM_MAX_CHUNK = 4


class PermSet(ctypes.Structure):
    """mzk_perm_set: one permutation set's columns and challenges, and its outputs."""
    _fields_ = [("values", ctypes.c_void_p * PERM_MAX_CHUNK), ("sigma", ctypes.c_void_p * PERM_MAX_CHUNK),
                ("delta_beta", (ctypes.c_uint64 * 4) * PERM_MAX_CHUNK), ("beta", ctypes.c_uint64 * 4),
                ("gamma", ctypes.c_uint64 * 4), ("ncols", ctypes.c_uint32), ("num", ctypes.c_void_p),
                ("den", ctypes.c_void_p)]


class PrefixJob(ctypes.Structure):
    """mzk_prefix_job: d_out = inclusive prefix products of d_a (n elements)."""
    _fields_ = [("d_a", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("n", ctypes.c_size_t)]


class KateJob(ctypes.Structure):
    """mzk_kate_job: d_q = quotient of d_a (n coefficients) by (X - b)."""
    _fields_ = [("d_a", ctypes.c_void_p), ("n", ctypes.c_size_t), ("b", ctypes.c_uint64 * 4),
                ("d_q", ctypes.c_void_p)]


class LogupArgs(ctypes.Structure):
    """mzk_logup_args: the log-derivative pass of one LogUp argument."""
    _fields_ = [("n", ctypes.c_size_t), ("blinding_factors", ctypes.c_size_t),
                ("selector", ctypes.c_void_p), ("multiplicities", ctypes.c_void_p), ("table", ctypes.c_void_p),
                ("inputs", ctypes.c_void_p), ("ninputs", ctypes.c_size_t), ("chunk_sizes", ctypes.c_void_p),
                ("nchunks", ctypes.c_size_t), ("beta", ctypes.c_uint64 * 4), ("blinding", ctypes.c_void_p),
                ("helpers", ctypes.c_void_p), ("aggregator", ctypes.c_void_p)]


class PlonkKey(ctypes.Structure):
    """mzk_plonk_key: the native prover's view of a proving key (device pointers)."""
    _fields_ = [("k", ctypes.c_uint32), ("g", ctypes.c_void_p), ("g_lagrange", ctypes.c_void_p),
                ("fixed_polys", ctypes.c_void_p * 5), ("fixed_cosets", ctypes.c_void_p * 5),
                ("sigma_values", ctypes.c_void_p * 5), ("sigma_polys", ctypes.c_void_p * 5),
                ("sigma_cosets", ctypes.c_void_p * 5),
                ("l0", ctypes.c_void_p), ("l_last", ctypes.c_void_p), ("l_active", ctypes.c_void_p),
                ("coset_x", ctypes.c_void_p), ("omega_powers", ctypes.c_void_p),
                ("transcript_repr", ctypes.c_uint64 * 4), ("g_lagrange_delta", ctypes.c_void_p)]


PLONK_PROOF_BYTES = 1360

_lib = None

# (name, argtypes, restype); every int-returning function is a status code
_P = ctypes.c_void_p
_S = ctypes.c_size_t
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_I = ctypes.c_int
_D = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "mzk_last_error": ([], ctypes.c_char_p),
    "mzk_version": ([], _I),
    "mzk_ctx_create": ([_I, ctypes.POINTER(_P)], _I),
    "mzk_ctx_destroy": ([_P], None),
    "mzk_ctx_stream": ([_P], _P),
    "mzk_ctx_synchronize": ([_P], _I),
    "mzk_ctx_mem_info": ([_P, ctypes.POINTER(_S), ctypes.POINTER(_S)], _I),
    "mzk_msm_best": ([_P, _P, _P, _S, _P], _I),
    "mzk_msm_specific": ([_P, _P, _P, _S, _P], _I),
    "mzk_msm_dev": ([_P, _P, _P, _S, _P], _I),
    "mzk_srs_create": ([_P, _P, _S, ctypes.POINTER(_P)], _I),
    "mzk_srs_generate": ([_P, _P, _S, ctypes.POINTER(_P)], _I),
    "mzk_srs_len": ([_P], _S),
    "mzk_srs_bases_dev": ([_P], _P),
    "mzk_srs_download": ([_P, _S, _S, _P], _I),
    "mzk_srs_destroy": ([_P], None),
    "mzk_msm_srs_dev": ([_P, _P, _S, _P, _S, _P], _I),
    "mzk_msm_srs_batch_dev": ([_P, _P, _P, _S, _P], _I),
    "mzk_g1_jacobian_to_affine": ([_P, _P], _I),
    "mzk_ntt_dev": ([_P, _P], _I),
    "mzk_ntt_batch_dev": ([_P, _P, _S], _I),
    "mzk_ntt_dev_noalloc": ([_P, _P], _I),
    "mzk_best_fft": ([_P, _P, _P, _U32], _I),
    "mzk_compute_twiddles": ([_P, _P, _U32, _P], _I),
    "mzk_best_fft_with_twiddles": ([_P, _P, _P, _U32], _I),
    "mzk_fft_coeff_to_extended": ([_P, _P, _P, _U32, _S], _I),
    "mzk_best_fft_dev": ([_P, _P, _P, _P, _U32, _U64], _I),
    "mzk_msm_parallel": ([_P, _P, _P, _S, _P], _I),
    "mzk_msm_serial": ([_P, _P, _P, _S, _P], _I),
    "mzk_domain_create": ([_P, _U32, _U32, ctypes.POINTER(_P)], _I),
    "mzk_domain_destroy": ([_P], None),
    "mzk_domain_info": ([_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32)], _I),
    "mzk_lagrange_to_coeff_dev": ([_P, _P], _I),
    "mzk_coeff_to_lagrange_dev": ([_P, _P], _I),
    "mzk_coeff_to_extended_dev": ([_P, _P, _P], _I),
    "mzk_extended_to_coeff_dev": ([_P, _P], _I),
    "mzk_divide_by_vanishing_poly_dev": ([_P, _P], _I),
    "mzk_lagrange_to_coeff": ([_P, _P], _I),
    "mzk_coeff_to_extended": ([_P, _P, _P], _I),
    "mzk_extended_to_coeff": ([_P, _P], _I),
    "mzk_fr_from_seed_dev": ([_P, _U64, _U64, _U64, _S, _P], _I),
    "mzk_diag_field_op": ([_P, _I, _I, _P, _P, _P, _S], _I),
    "mzk_diag_g1_op": ([_P, _I, _P, _P, _P, _S], _I),
    "mzk_diag_bench": ([_P, _I, _I, _I, _I, _D, _D], _I),
    "mzk_prof_enable": ([_P, _I], _I),
    "mzk_ctx_set_graphs": ([_P, _I], _I),
    "mzk_ctx_graph_stats": ([_P, _P], _I),
    "mzk_prof_read": ([_P, ctypes.c_char_p, _S], _I),
    "mzk_msm_last_plan": ([_P, _P], _I),
    "mzk_srs_from_scalars_dev": ([_P, _P, _S, ctypes.POINTER(_P)], _I),
    "mzk_srs_to_lagrange": ([_P, _P, _U32, ctypes.POINTER(_P)], _I),
    "mzk_srs_suffix_sum": ([_P, _P, ctypes.POINTER(_P)], _I),
    "mzk_plonk_perm_products_dev": ([_P, _P, _S, _P, _S], _I),
    "mzk_tune_msm": ([_P, _P, _U32, _U32, _U32], _I),
    "mzk_tune_ntt": ([_P, _U32, _U32, _U32], _I),
    "mzk_tune_export": ([_P, _S, ctypes.POINTER(_S)], _I),
    "mzk_tune_import": ([ctypes.c_char_p], _I),
    "mzk_tune_reset": ([], None),
    "mzk_srs_read_points": ([_P, _P, _S, _I, ctypes.POINTER(_P)], _I),
    "mzk_srs_write_points": ([_P, _P, _S, _S, _I, _P], _I),
    "mzk_fr_vec_op": ([_P, _I, _P, _P, _P, _S], _I),
    "mzk_fr_vec_scalar": ([_P, _I, _P, _P, _P, _S], _I),
    "mzk_fr_batch_invert": ([_P, _P, _S], _I),
    "mzk_fr_prefix_product": ([_P, _P, _P, _S], _I),
    "mzk_fr_eval": ([_P, _P, _S, _P, _P], _I),
    "mzk_fr_eval_batch": ([_P, _P, _P, _P, _S, _P], _I),
    "mzk_plonk_numerator_dev": ([_P, _P, _P], _I),
    "mzk_plonk_prove": ([_P, _P, _S, _P, _P, _P, _P], _I),
    "mzk_fr_lincomb_dev": ([_P, _P, _P, _P, _S, _P, _S], _I),
    "mzk_fr_kate_div": ([_P, _P, _S, _P, _P], _I),
    "mzk_fr_kate_div_batch": ([_P, _P, _S], _I),
    "mzk_fr_prefix_product_batch": ([_P, _P, _S], _I),
    "mzk_fr_powers": ([_P, _P, _S, _P], _I),
    "mzk_fr_to_delta_dev": ([_P, _P, _P, _S, _I], _I),
    "mzk_fr_prefix_sum": ([_P, _P, _P, _S], _I),
    "mzk_logup_multiplicities_dev": ([_P, _P, _P, _S, _P, _S, _S, _P, _P], _I),
    "mzk_logup_logderivative_dev": ([_P, _P], _I),
}


def load_library(path: str = LIB_PATH):
    """Loads libmzk_gpu.so and binds every declared symbol (raises if any is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MzkError(f"{path} is not built; run midnight-zk_amd/build.py")
    # torch ships its own libamdhip64.so.7.  Loading torch first makes this library
    # bind to that same HIP runtime; the other order puts two runtimes in the
    # process and the second one finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(L, name)  # AttributeError if the export is missing
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        msg = load_library().mzk_last_error().decode(errors="replace")
        raise MzkError(f"mzk status {rc}: {msg}")


def _np_ptr(a: np.ndarray):
    if a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"]:
        raise MzkError("expected a C-contiguous uint64 array")
    return a.ctypes.data_as(ctypes.c_void_p)


def _byte_ptr(a: np.ndarray):
    if a.dtype != np.uint8 or not a.flags["C_CONTIGUOUS"]:
        raise MzkError("expected a C-contiguous uint8 array")
    return a.ctypes.data_as(ctypes.c_void_p)


def _t_ptr(t) -> int:
    if not t.is_cuda or not t.is_contiguous():
        raise MzkError("expected a contiguous device tensor")
    return t.data_ptr()


@contextlib.contextmanager
def _order(ctx):
    """Orders work on the context's stream with the caller's current torch stream: the
    context stream first waits for the torch stream (the tensors' producers), and the
    torch stream afterwards waits for the context stream (their consumers), so a call
    orders like a torch op on the current stream.  Nothing is added when the current
    torch stream is the context's own, or when torch has not touched the GPU."""
    torch = sys.modules.get("torch")
    cur = None
    if torch is not None and torch.cuda.is_initialized():
        c = torch.cuda.current_stream()
        if c.cuda_stream != ctx.stream:
            cur = c
            ev = torch.cuda.Event()
            ev.record(c)
            ctx.torch_stream().wait_event(ev)
    yield
    if cur is not None:
        ev = torch.cuda.Event()
        ev.record(ctx.torch_stream())
        cur.wait_event(ev)


def _ordered(ctx_of):
    """Decorator: the call runs inside _order(ctx_of(self))."""
    def deco(fn):
        def run(self, *a, **kw):
            with _order(ctx_of(self)):
                return fn(self, *a, **kw)
        run.__name__, run.__doc__ = fn.__name__, fn.__doc__
        return run
    return deco


def _self_ctx(obj):
    return obj


def _owner_ctx(obj):
    return obj.ctx


class Context:
    """One device context (stream + scratch).  Mirrors no reference type: the
    reference runs on the rayon pool; this is the device it runs on."""

    def __init__(self, device: int = 0):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.mzk_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            load_library().mzk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return load_library().mzk_ctx_stream(self.h) or 0

    def torch_stream(self):
        """The context's stream as a torch stream (for ordering torch work with it)."""
        if getattr(self, "_tstream", None) is None:
            import torch
            self._tstream = torch.cuda.ExternalStream(self.stream, device=self.device)
        return self._tstream

    def synchronize(self):
        _check(load_library().mzk_ctx_synchronize(self.h))

    def mem_info(self) -> tuple[int, int]:
        """(free, total) bytes of the device's memory (hipMemGetInfo)."""
        f, t = _S(), _S()
        _check(load_library().mzk_ctx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def set_graphs(self, on: bool = True):
        """Replay cached HIP graphs of the MSM / NTT launch sequences (default on)."""
        _check(load_library().mzk_ctx_set_graphs(self.h, 1 if on else 0))

    def graph_stats(self) -> dict:
        out = np.zeros(3, dtype=np.uint64)
        _check(load_library().mzk_ctx_graph_stats(self.h, out.ctypes.data))
        return {"captured": int(out[0]), "replayed": int(out[1]), "held": int(out[2])}

    def profile(self, on: bool = True):
        """Per-kernel HIP-event timing on this context's stream (off by default)."""
        _check(load_library().mzk_prof_enable(self.h, 1 if on else 0))

    def profile_read(self) -> dict:
        """{kernel name: (total ms, launches)} since the last read; clears the totals."""
        buf = ctypes.create_string_buffer(1 << 16)
        _check(load_library().mzk_prof_read(self.h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, cnt = line.split()
            out[name] = (float(ms), int(cnt))
        return out

    def msm_last_plan(self) -> dict:
        """Plan of the last MSM on this context (window bits, digits, ...)."""
        out = np.zeros(6, dtype=np.uint64)
        _check(load_library().mzk_msm_last_plan(self.h, _np_ptr(out)))
        keys = ("c", "windows", "bucket_sets", "refs_per_thread", "digits", "n")
        return {k: int(v) for k, v in zip(keys, out)}

    # ------------------------------------------------------------------ MSM
    def msm_best(self, coeffs: np.ndarray, bases: np.ndarray) -> np.ndarray:
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64).reshape(-1, 4)
        bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
        if coeffs.shape[0] != bases.shape[0]:
            raise MzkError("msm_best: coeffs and bases differ in length")
        out = np.zeros(18, dtype=np.uint64)
        _check(load_library().mzk_msm_best(self.h, _np_ptr(coeffs), _np_ptr(bases), coeffs.shape[0], _np_ptr(out)))
        return out

    def msm_specific(self, coeffs: np.ndarray, bases: np.ndarray) -> np.ndarray:
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64).reshape(-1, 4)
        bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
        if coeffs.shape[0] != bases.shape[0]:
            raise MzkError("msm_specific: coeffs and bases differ in length")
        out = np.zeros(18, dtype=np.uint64)
        _check(load_library().mzk_msm_specific(self.h, _np_ptr(coeffs), _np_ptr(bases), coeffs.shape[0], _np_ptr(out)))
        return out

    def msm_parallel(self, coeffs: np.ndarray, bases: np.ndarray) -> np.ndarray:
        coeffs, bases = self._msm_args(coeffs, bases, "msm_parallel")
        out = np.zeros(18, dtype=np.uint64)
        _check(load_library().mzk_msm_parallel(self.h, _np_ptr(coeffs), _np_ptr(bases), coeffs.shape[0], _np_ptr(out)))
        return out

    def msm_serial(self, coeffs: np.ndarray, bases: np.ndarray, acc: np.ndarray) -> np.ndarray:
        """acc (18 Jacobian limbs) += sum coeffs[i] bases[i]; returns the updated acc."""
        coeffs, bases = self._msm_args(coeffs, bases, "msm_serial")
        acc = np.ascontiguousarray(acc, dtype=np.uint64).reshape(18).copy()
        _check(load_library().mzk_msm_serial(self.h, _np_ptr(coeffs), _np_ptr(bases), coeffs.shape[0], _np_ptr(acc)))
        return acc

    @staticmethod
    def _msm_args(coeffs, bases, who):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64).reshape(-1, 4)
        bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
        if coeffs.shape[0] != bases.shape[0]:
            raise MzkError(f"{who}: coeffs and bases differ in length")
        return coeffs, bases

    @_ordered(_self_ctx)
    def msm_dev(self, d_coeffs, d_bases, n: int) -> np.ndarray:
        out = np.zeros(18, dtype=np.uint64)
        _check(load_library().mzk_msm_dev(self.h, _t_ptr(d_coeffs), _t_ptr(d_bases), n, _np_ptr(out)))
        return out

    # ------------------------------------------------------------------ NTT
    def best_fft(self, a: np.ndarray, omega: np.ndarray, log_n: int) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4).copy()
        if a.shape[0] != (1 << log_n):
            raise MzkError("best_fft: len(a) != 2^log_n")
        om = np.ascontiguousarray(omega, dtype=np.uint64).reshape(4)
        _check(load_library().mzk_best_fft(self.h, _np_ptr(a), _np_ptr(om), log_n))
        return a

    def compute_twiddles(self, omega: np.ndarray, log_n: int) -> np.ndarray:
        out = np.zeros((1 << max(0, log_n - 1), 4), dtype=np.uint64)
        om = np.ascontiguousarray(omega, dtype=np.uint64).reshape(4)
        _check(load_library().mzk_compute_twiddles(self.h, _np_ptr(om), log_n, _np_ptr(out)))
        return out

    def best_fft_with_twiddles(self, a: np.ndarray, twiddles: np.ndarray, log_n: int) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4).copy()
        if a.shape[0] != (1 << log_n):
            raise MzkError("best_fft_with_twiddles: len(a) != 2^log_n")
        tw = np.ascontiguousarray(twiddles, dtype=np.uint64).reshape(-1, 4)
        if tw.shape[0] < max(1, (1 << log_n) // 2):
            raise MzkError("best_fft_with_twiddles: twiddle table shorter than 2^(log_n-1)")
        _check(load_library().mzk_best_fft_with_twiddles(self.h, _np_ptr(a), _np_ptr(tw), log_n))
        return a

    def fft_coeff_to_extended(self, a: np.ndarray, twiddles: np.ndarray, log_n: int, n_real: int) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4).copy()
        if a.shape[0] != (1 << log_n):
            raise MzkError("fft_coeff_to_extended: len(a) != 2^log_n")
        tw = np.ascontiguousarray(twiddles, dtype=np.uint64).reshape(-1, 4)
        if tw.shape[0] < max(1, (1 << log_n) // 2):
            raise MzkError("fft_coeff_to_extended: twiddle table shorter than 2^(log_n-1)")
        _check(load_library().mzk_fft_coeff_to_extended(self.h, _np_ptr(a), _np_ptr(tw), log_n, n_real))
        return a

    @_ordered(_self_ctx)
    def best_fft_dev(self, d_in, d_out, omega: np.ndarray, log_n: int, n_real: int | None = None):
        om = np.ascontiguousarray(omega, dtype=np.uint64).reshape(4)
        n_real = (1 << log_n) if n_real is None else n_real
        _check(load_library().mzk_best_fft_dev(self.h, _t_ptr(d_in), _t_ptr(d_out), _np_ptr(om), log_n, n_real))

    @_ordered(_self_ctx)
    def ntt(self, job: "NttJob"):
        _check(load_library().mzk_ntt_dev(self.h, ctypes.byref(job)))

    @_ordered(_self_ctx)
    def ntt_noalloc(self, job: "NttJob"):
        """mzk_ntt_dev_noalloc: the same transform from the library's fixed per-device
        array (log_n <= NTT_NOALLOC_MAX_LOG), no device allocation."""
        _check(load_library().mzk_ntt_dev_noalloc(self.h, ctypes.byref(job)))

    @_ordered(_self_ctx)
    def ntt_batch(self, jobs):
        jobs = list(jobs)
        arr = (NttJob * max(1, len(jobs)))(*jobs)
        _check(load_library().mzk_ntt_batch_dev(self.h, ctypes.cast(arr, ctypes.c_void_p), len(jobs)))

    def domain(self, j: int, k: int) -> "EvaluationDomain":
        return EvaluationDomain(self, j, k)

    def srs_generate(self, tau_mont: np.ndarray, n: int) -> "ResidentSrs":
        return ResidentSrs.generate(self, tau_mont, n)

    def srs_upload(self, bases: np.ndarray) -> "ResidentSrs":
        return ResidentSrs.upload(self, bases)

    # ------------------------------------------------------------------ LogUp
    @_ordered(_self_ctx)
    def logup_multiplicities(self, d_selector, d_inputs, d_table, usable_rows: int, blinding: np.ndarray, d_out):
        """compute_multiplicities (logup/prover.rs:411-477) into d_out; blinding: host
        (n - usable_rows, 4) Montgomery rows."""
        n = d_table.shape[0]
        ins = (ctypes.c_void_p * max(1, len(d_inputs)))(*[_t_ptr(d) for d in d_inputs])
        bl = np.ascontiguousarray(blinding, dtype=np.uint64).reshape(-1, 4)
        if bl.shape[0] != n - usable_rows:
            raise MzkError("logup_multiplicities: one blinding value per row from usable_rows on")
        _check(load_library().mzk_logup_multiplicities_dev(
            self.h, _t_ptr(d_selector), ctypes.cast(ins, ctypes.c_void_p), len(d_inputs), _t_ptr(d_table), n,
            usable_rows, _np_ptr(bl) if bl.shape[0] else None, _t_ptr(d_out)))

    @_ordered(_self_ctx)
    def logup_logderivative(self, d_selector, d_m, d_table, d_inputs, chunk_sizes, beta_mont: np.ndarray,
                            blinding: np.ndarray, d_helpers, d_aggregator):
        """compute_logderivative (logup/prover.rs:189-310): helper columns and aggregator."""
        a = LogupArgs()
        bl = np.ascontiguousarray(blinding, dtype=np.uint64).reshape(-1, 4)
        ins = (ctypes.c_void_p * max(1, len(d_inputs)))(*[_t_ptr(d) for d in d_inputs])
        hs = (ctypes.c_void_p * max(1, len(d_helpers)))(*[_t_ptr(d) for d in d_helpers])
        cs = np.asarray(chunk_sizes, dtype=np.uint32)
        a.n = d_table.shape[0]
        a.blinding_factors = bl.shape[0]
        a.selector, a.multiplicities, a.table = _t_ptr(d_selector), _t_ptr(d_m), _t_ptr(d_table)
        a.inputs, a.ninputs = ctypes.cast(ins, ctypes.c_void_p), len(d_inputs)
        a.chunk_sizes, a.nchunks = cs.ctypes.data, len(cs)
        a.beta[:] = [int(w) for w in np.asarray(beta_mont, dtype=np.uint64).reshape(4)]
        a.blinding = bl.ctypes.data if bl.shape[0] else None
        a.helpers, a.aggregator = ctypes.cast(hs, ctypes.c_void_p), _t_ptr(d_aggregator)
        _check(load_library().mzk_logup_logderivative_dev(self.h, ctypes.byref(a)))

    @_ordered(_self_ctx)
    def fr_prefix_sum(self, d_a, d_out, n: int | None = None):
        n = d_a.shape[0] if n is None else n
        _check(load_library().mzk_fr_prefix_sum(self.h, _t_ptr(d_a), _t_ptr(d_out), n))

    # ------------------------------------------------------------------ inputs
    @_ordered(_self_ctx)
    def fr_from_seed_dev(self, seed: int, stream: int, n: int, out, offset: int = 0):
        _check(load_library().mzk_fr_from_seed_dev(self.h, seed, stream, offset, n, _t_ptr(out)))
        return out


class EvaluationDomain:
    """proofs/src/poly/domain.rs:34-167; transforms on device tensors of shape
    (len, 4) uint64 (viewed as int64 by torch)."""

    def __init__(self, ctx: Context, j: int, k: int):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(load_library().mzk_domain_create(ctx.h, j, k, ctypes.byref(h)))
        self.h = h
        kk, ek = _U32(), _U32()
        _check(load_library().mzk_domain_info(self.h, ctypes.byref(kk), ctypes.byref(ek)))
        self.k = kk.value
        self.extended_k = ek.value

    def close(self):
        if self.h:
            load_library().mzk_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @_ordered(_owner_ctx)
    def lagrange_to_coeff(self, d_a):
        _check(load_library().mzk_lagrange_to_coeff_dev(self.h, _t_ptr(d_a)))
        return d_a

    @_ordered(_owner_ctx)
    def coeff_to_lagrange(self, d_a):
        _check(load_library().mzk_coeff_to_lagrange_dev(self.h, _t_ptr(d_a)))
        return d_a

    @_ordered(_owner_ctx)
    def coeff_to_extended(self, d_in, d_out):
        _check(load_library().mzk_coeff_to_extended_dev(self.h, _t_ptr(d_in), _t_ptr(d_out)))
        return d_out

    @_ordered(_owner_ctx)
    def extended_to_coeff(self, d_a):
        _check(load_library().mzk_extended_to_coeff_dev(self.h, _t_ptr(d_a)))
        return d_a

    @_ordered(_owner_ctx)
    def divide_by_vanishing_poly(self, d_a):
        _check(load_library().mzk_divide_by_vanishing_poly_dev(self.h, _t_ptr(d_a)))
        return d_a

    # host-pointer forms (copy in / out)
    def lagrange_to_coeff_host(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4).copy()
        _check(load_library().mzk_lagrange_to_coeff(self.h, _np_ptr(a)))
        return a

    def coeff_to_extended_host(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
        out = np.zeros((1 << self.extended_k, 4), dtype=np.uint64)
        _check(load_library().mzk_coeff_to_extended(self.h, _np_ptr(a), _np_ptr(out)))
        return out

    def extended_to_coeff_host(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4).copy()
        _check(load_library().mzk_extended_to_coeff(self.h, _np_ptr(a)))
        return a


class ResidentSrs:
    """Affine bases resident in HBM (proofs/src/poly/kzg/params.rs:22-37)."""

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h
        self.n = load_library().mzk_srs_len(self.h)

    @classmethod
    def generate(cls, ctx: Context, tau_mont: np.ndarray, n: int) -> "ResidentSrs":
        h = ctypes.c_void_p()
        tau = np.ascontiguousarray(tau_mont, dtype=np.uint64).reshape(4)
        _check(load_library().mzk_srs_generate(ctx.h, _np_ptr(tau), n, ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def upload(cls, ctx: Context, bases: np.ndarray) -> "ResidentSrs":
        h = ctypes.c_void_p()
        b = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
        _check(load_library().mzk_srs_create(ctx.h, _np_ptr(b), b.shape[0], ctypes.byref(h)))
        return cls(ctx, h)

    def download(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        count = self.n - offset if count is None else count
        out = np.zeros((count, 12), dtype=np.uint64)
        _check(load_library().mzk_srs_download(self.h, offset, count, _np_ptr(out)))
        return out

    def to_lagrange(self, k: int) -> "ResidentSrs":
        """g_to_lagrange (utils/arithmetic.rs:34-54): Lagrange bases of the first 2^k bases."""
        h = ctypes.c_void_p()
        _check(load_library().mzk_srs_to_lagrange(self.ctx.h, self.h, k, ctypes.byref(h)))
        return ResidentSrs(self.ctx, h)

    def suffix_sum(self) -> "ResidentSrs":
        """suffix_sum (kzg/params.rs:47-57): the delta bases of these bases."""
        h = ctypes.c_void_p()
        _check(load_library().mzk_srs_suffix_sum(self.ctx.h, self.h, ctypes.byref(h)))
        return ResidentSrs(self.ctx, h)

    @classmethod
    def read_points(cls, ctx: Context, data, count: int, fmt: int) -> "ResidentSrs":
        """count encoded G1 points (48 B compressed for SERDE_PROCESSED, else 96 B) decoded
        and checked on the device (ParamsKZG::read_custom's point arrays,
        kzg/params.rs:286-336)."""
        buf = np.frombuffer(data, dtype=np.uint8)
        need = count * point_bytes(fmt)
        if buf.size < need:
            raise ValueError(f"{need} bytes of points needed, {buf.size} given")
        buf = np.ascontiguousarray(buf[:need])
        h = ctypes.c_void_p()
        _check(load_library().mzk_srs_read_points(ctx.h, _byte_ptr(buf), count, fmt, ctypes.byref(h)))
        return cls(ctx, h)

    def write_points(self, fmt: int, offset: int = 0, count: int | None = None) -> bytes:
        """The bases encoded as in ParamsKZG::write_custom (kzg/params.rs:259-281)."""
        count = self.n - offset if count is None else count
        out = np.zeros(count * point_bytes(fmt), dtype=np.uint8)
        _check(load_library().mzk_srs_write_points(self.ctx.h, self.h, offset, count, fmt, _byte_ptr(out)))
        return out.tobytes()

    def bases_ptr(self) -> int:
        return load_library().mzk_srs_bases_dev(self.h)

    def commit_batch(self, jobs, ctx: "Context | None" = None) -> np.ndarray:
        """Many commitments in one pipeline.  jobs: iterable of (device coeffs, n, offset);
        returns an (len(jobs), 18) array of Jacobian results.  ctx: the context (stream
        and scratch) to run on; the bases are read-only, so any context of the same
        device may use them (default: the one that built them)."""
        ctx = ctx or self.ctx
        jobs = list(jobs)
        arr = (MsmJob * max(1, len(jobs)))()
        for i, (d, n, off) in enumerate(jobs):
            arr[i].d_coeffs = _t_ptr(d) if n else None
            arr[i].n = n
            arr[i].offset = off
        out = np.zeros((max(1, len(jobs)), 18), dtype=np.uint64)
        with _order(ctx):
            _check(load_library().mzk_msm_srs_batch_dev(ctx.h, self.h, ctypes.cast(arr, ctypes.c_void_p),
                                                         len(jobs), _np_ptr(out)))
        return out[:len(jobs)]

    @_ordered(_owner_ctx)
    def commit(self, d_coeffs, n: int | None = None, offset: int = 0) -> np.ndarray:
        """KZG commit of device coefficients against bases[offset:offset+n]."""
        n = d_coeffs.shape[0] if n is None else n
        out = np.zeros(18, dtype=np.uint64)
        _check(load_library().mzk_msm_srs_dev(self.ctx.h, self.h, offset, _t_ptr(d_coeffs), n, _np_ptr(out)))
        return out

    def close(self):
        if self.h:
            load_library().mzk_srs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------- autotuner
def tune_msm(ctx: Context, srs: "ResidentSrs", log_lo: int, log_hi: int, reps: int = 3) -> None:
    """Time the MSM plan candidates at 2^log_lo..2^log_hi against srs; keep the fastest."""
    _check(load_library().mzk_tune_msm(ctx.h, srs.h, log_lo, log_hi, reps))


def tune_ntt(ctx: Context, log_lo: int, log_hi: int, reps: int = 3) -> None:
    """Time the NTT split / workgroup candidates at 2^log_lo..2^log_hi; keep the fastest."""
    _check(load_library().mzk_tune_ntt(ctx.h, log_lo, log_hi, reps))


def tune_export() -> str:
    L = load_library()
    need = ctypes.c_size_t(0)
    _check(L.mzk_tune_export(None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    _check(L.mzk_tune_export(buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


def tune_import(text: str) -> None:
    _check(load_library().mzk_tune_import(text.encode()))


def tune_reset() -> None:
    load_library().mzk_tune_reset()


# SerdeFormat (proofs/src/utils/helpers.rs:16-28), in the reference's order
SERDE_PROCESSED, SERDE_RAW_BYTES, SERDE_RAW_BYTES_UNCHECKED = 0, 1, 2


def point_bytes(fmt: int, g2: bool = False) -> int:
    """Encoded size of one G1 (or G2) point: compressed for Processed, else uncompressed."""
    if fmt not in (SERDE_PROCESSED, SERDE_RAW_BYTES, SERDE_RAW_BYTES_UNCHECKED):
        raise ValueError(f"unknown SerdeFormat {fmt}")
    base = 96 if g2 else 48
    return base if fmt == SERDE_PROCESSED else 2 * base


class KzgParams:
    """ParamsKZG (proofs/src/poly/kzg/params.rs:22-37) with its four G1 bases resident in
    HBM: g, g_lagrange, and the delta bases suffix_sum(g_lagrange) and
    suffix_sum(g_lagrange_delta) that read_custom derives (params.rs:338-340).  The G2
    points are kept as the bytes read (the prover's MSMs never use them)."""

    def __init__(self, k: int, g: ResidentSrs, g_lagrange: ResidentSrs, g2: bytes, s_g2: bytes, g2_format: int):
        self.k = k
        self.g = g
        self.g_lagrange = g_lagrange
        self.g_lagrange_delta = g_lagrange.suffix_sum()
        self.g_lagrange_double_delta = self.g_lagrange_delta.suffix_sum()
        self.g2, self.s_g2, self.g2_format = g2, s_g2, g2_format

    @classmethod
    def read_custom(cls, ctx: Context, data, fmt: int) -> "KzgParams":
        """ParamsKZG::read_custom (params.rs:284-349): k as u32 LE, 2^k points g, 2^k points
        g_lagrange, then g2 and s_g2.  Raises MzkError on a point that fails its checks
        and ValueError on a short buffer."""
        mv = memoryview(data).cast("B")
        if len(mv) < 4:
            raise ValueError("truncated parameters: no k")
        k = int.from_bytes(mv[:4], "little")
        if k > 30:
            raise ValueError(f"k = {k} out of range")
        n = 1 << k
        pb, gb = point_bytes(fmt), point_bytes(fmt, g2=True)
        need = 4 + 2 * n * pb + 2 * gb
        if len(mv) < need:
            raise ValueError(f"truncated parameters: {need} bytes needed, {len(mv)} given")
        g = ResidentSrs.read_points(ctx, mv[4:4 + n * pb], n, fmt)
        gl = ResidentSrs.read_points(ctx, mv[4 + n * pb:4 + 2 * n * pb], n, fmt)
        o = 4 + 2 * n * pb
        return cls(k, g, gl, bytes(mv[o:o + gb]), bytes(mv[o + gb:o + 2 * gb]), fmt)

    def write_custom(self, fmt: int) -> bytes:
        """ParamsKZG::write_custom (params.rs:259-281).  The G2 points are written as read,
        so fmt must select the same G2 encoding (compressed or not) they came in."""
        if (fmt == SERDE_PROCESSED) != (self.g2_format == SERDE_PROCESSED):
            raise NotImplementedError("re-encoding G2 points between compressed and uncompressed forms")
        return (self.k.to_bytes(4, "little") + self.g.write_points(fmt) + self.g_lagrange.write_points(fmt)
                + self.g2 + self.s_g2)


def g1_to_affine(jac: np.ndarray) -> np.ndarray:
    """G1Projective::to_affine: 18-limb Jacobian -> 12-limb affine (identity = zeros)."""
    jac = np.ascontiguousarray(jac, dtype=np.uint64).reshape(18)
    out = np.zeros(12, dtype=np.uint64)
    _check(load_library().mzk_g1_jacobian_to_affine(_np_ptr(jac), _np_ptr(out)))
    return out


def diag_field_op(ctx: Context, field: int, op: int, d_a, d_b, d_out, n: int):
    _check(load_library().mzk_diag_field_op(ctx.h, field, op, _t_ptr(d_a), _t_ptr(d_b), _t_ptr(d_out), n))


def diag_g1_op(ctx: Context, op: int, d_a, d_b, d_out, n: int):
    _check(load_library().mzk_diag_g1_op(ctx.h, op, _t_ptr(d_a), _t_ptr(d_b), _t_ptr(d_out), n))


def diag_bench(ctx: Context, which: int, blocks: int, threads: int, iters: int):
    ms, ops = ctypes.c_double(), ctypes.c_double()
    _check(load_library().mzk_diag_bench(ctx.h, which, blocks, threads, iters, ctypes.byref(ms), ctypes.byref(ops)))
    return ms.value, ops.value
