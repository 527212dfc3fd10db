"""The library's native prover (mzk_plonk_prove, a group of proofs in lockstep) makes
the same bytes as the Python prover on the GPU backend for every proof of the group,
and its proofs pass the trapdoor verifier (GPU)."""
import time

import pytest

from tests import plonk_verifier as V
from tests.helpers import pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk


@pytest.fixture(scope="module")
def gpu():
    return plonk.GpuBackend()


@pytest.mark.parametrize("k,group", [(6, 3), (8, 3), (12, 2), (16, 2), (18, 2)])
def test_native_group_equals_python(gpu, k, group):
    print(f"\n[native prover k={k} group={group}]", flush=True)
    params = plonk.Params(gpu, k, 0xABCDEF + k)
    circ = plonk.SyntheticCircuit(k, seed=7 + k)
    pk = plonk.keygen(gpu, params, circ)
    wits = [circ.witness(gpu, 1000 + i) for i in range(group)]
    seeds = [50 + i for i in range(group)]
    ph = [0.0] * 8
    plonk.create_proofs_native(gpu, params, pk, wits, seeds)       # warm-up
    t0 = time.perf_counter()
    native = plonk.create_proofs_native(gpu, params, pk, wits, seeds, ph)
    t1 = time.perf_counter()
    print(f"  native group {1e3 * (t1 - t0):.1f} ms, phases {[round(x, 2) for x in ph]}", flush=True)
    for i in range(group):
        py = plonk.create_proof(gpu, params, pk, wits[i], seeds[i])
        assert native[i] == py, i
    if k <= 12:
        fixed = [V.affine_from_limbs(c) for c in pk.fixed_commitments]
        perm = [V.affine_from_limbs(c) for c in pk.permutation_commitments]
        assert V.verify(native[0], k, pk.transcript_repr, fixed, perm, params.tau)


def test_native_rejects_bad_arguments(gpu):
    import ctypes
    mzk = gpu.mzk
    key = mzk.PlonkKey()
    key.k = 30
    out = (ctypes.c_uint8 * 1360)()
    seeds = (ctypes.c_uint64 * 1)(1)
    cols = (ctypes.c_void_p * 5)()
    rc = gpu._L.mzk_plonk_prove(gpu.ctx.h, ctypes.addressof(key), 1, ctypes.cast(cols, ctypes.c_void_p),
                                ctypes.cast(seeds, ctypes.c_void_p), ctypes.cast(out, ctypes.c_void_p), None)
    assert rc == 1


def test_native_concurrent_calls_on_one_context(gpu):
    """Two threads proving on one context at once: each call holds the context's prover
    lock for its whole run, so both get the proofs a lone call gives."""
    import threading
    k = 8
    params = plonk.Params(gpu, k, 0xC0C0 + k)
    circ = plonk.SyntheticCircuit(k, seed=3 + k)
    pk = plonk.keygen(gpu, params, circ)
    wits = [circ.witness(gpu, 2000 + i) for i in range(4)]
    alone = [plonk.create_proofs_native(gpu, params, pk, [w], [9 + i])[0] for i, w in enumerate(wits)]
    got = [None, None]

    def run(t):
        got[t] = plonk.create_proofs_native(gpu, params, pk, wits[2 * t:2 * t + 2], [9 + 2 * t, 10 + 2 * t])

    th = [threading.Thread(target=run, args=(t,)) for t in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert got[0] + got[1] == alone
