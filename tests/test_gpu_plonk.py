"""The PLONK prover on the GPU backend: keys and proofs are byte-identical to the
CPU backend's (the oracle's C restatement running the same prover), and proofs pass
the trapdoor verifier up to the k = 18 configuration (GPU)."""
import time

import numpy as np
import pytest

from tests import plonk_verifier as V
from tests.cpu_backend import CpuBackend
from tests.helpers import from_dev, pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk


@pytest.fixture(scope="module")
def gpu():
    b = plonk.GpuBackend()
    yield b


@pytest.fixture(scope="module")
def cpu():
    return CpuBackend(16)


def verify(params, pk, proof):
    fixed = [V.affine_from_limbs(c) for c in pk.fixed_commitments]
    perm = [V.affine_from_limbs(c) for c in pk.permutation_commitments]
    return V.verify(proof, pk.k, pk.transcript_repr, fixed, perm, params.tau)


@pytest.mark.parametrize("k", [6, 8, 12, 16])
def test_gpu_proof_identical_to_cpu(gpu, cpu, k):
    print(f"\n[plonk identity k={k}]", flush=True)
    tau = 0xC0FFEE + k
    pg, pc = plonk.Params(gpu, k, tau), plonk.Params(cpu, k, tau)
    assert np.array_equal(pg.g_lagrange.download(), pc.g_lagrange.bases)
    circ = plonk.SyntheticCircuit(k, seed=41)
    kg, kc = plonk.keygen(gpu, pg, circ), plonk.keygen(cpu, pc, circ)
    assert kg.transcript_repr == kc.transcript_repr
    wg, wc = circ.witness(gpu), circ.witness(cpu)
    gpu.sync()
    for a, b in zip(wg, wc):
        assert np.array_equal(from_dev(a).reshape(-1, 4), b)
    t0 = time.perf_counter()
    proof_g = plonk.create_proof(gpu, pg, kg, wg, rng_seed=1234)
    t1 = time.perf_counter()
    proof_c = plonk.create_proof(cpu, pc, kc, wc, rng_seed=1234)
    t2 = time.perf_counter()
    print(f"  gpu prove {1e3 * (t1 - t0):.1f} ms, cpu prove {1e3 * (t2 - t1):.1f} ms", flush=True)
    assert proof_g == proof_c
    assert verify(pg, kg, proof_g)


def test_gpu_k18_proof_verifies(gpu):
    k = 18
    print("\n[plonk k=18 gpu]", flush=True)
    params = plonk.Params(gpu, k, 0x5EED18)
    circ = plonk.SyntheticCircuit(k, seed=18)
    pk = plonk.keygen(gpu, params, circ)
    wit = circ.witness(gpu)
    plonk.create_proof(gpu, params, pk, wit, rng_seed=7)       # warm-up
    gpu.sync()
    t0 = time.perf_counter()
    proof = plonk.create_proof(gpu, params, pk, wit, rng_seed=7)
    print(f"  prove {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    assert verify(params, pk, proof)
    # the proof is deterministic in its seeds
    assert plonk.create_proof(gpu, params, pk, wit, rng_seed=7) == proof
