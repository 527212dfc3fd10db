"""The PLONK prover's host logic on the CPU backend (oracle C restatement): proofs
of the synthetic circuit pass the trapdoor verifier, broken witnesses and tampered
proofs are rejected, and the circuit's witness satisfies its gate and copies."""
import hashlib

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests import plonk_verifier as V
from tests.cpu_backend import CpuBackend
from tests.helpers import pkg

plonk = pkg.plonk


@pytest.fixture(scope="module")
def be():
    return CpuBackend(8)


def setup(be, k, seed, tau):
    params = plonk.Params(be, k, tau)
    circ = plonk.SyntheticCircuit(k, seed)
    pk = plonk.keygen(be, params, circ)
    return params, circ, pk


def verify(params, pk, proof):
    fixed = [V.affine_from_limbs(c) for c in pk.fixed_commitments]
    perm = [V.affine_from_limbs(c) for c in pk.permutation_commitments]
    return V.verify(proof, pk.k, pk.transcript_repr, fixed, perm, params.tau)


def test_constants_match_the_oracle():
    assert plonk.R_MOD == O.R_MOD and plonk.P_MOD == O.P_MOD
    assert plonk.ZETA == O.FR_ZETA and plonk.DELTA == O.FR_DELTA
    assert plonk.omega_for(10) == O.root_of_unity(10)


def test_compression_matches_the_oracle():
    for s in (1, 2, 12345, O.R_MOD - 1):
        P = O.g1_mul(O.G1_GEN, s)
        limbs = O.g1_affine_mont_limbs(P)
        assert plonk.g1_compress(limbs) == O.g1_to_compressed(P)
    assert plonk.g1_compress(np.zeros(12, dtype=np.uint64)) == O.g1_to_compressed(None)


def test_transcript_is_keyed_blake2b():
    """Absorb = 0x01 || repr, squeeze = 0x00 then the digest (implementors.rs:13-29),
    checked against hashlib directly."""
    t = plonk.Transcript()
    t.write_fr(5)
    c1 = t.squeeze()
    h = hashlib.blake2b(digest_size=64, key=b"Domain separator for transcript")
    h.update(b"\x01" + (5).to_bytes(32, "little"))
    h.update(b"\x00")
    assert c1 == int.from_bytes(h.digest(), "little") % O.R_MOD
    h.update(b"\x00")
    assert t.squeeze() == int.from_bytes(h.digest(), "little") % O.R_MOD
    assert bytes(t.proof) == (5).to_bytes(32, "little")


@pytest.mark.parametrize("k", [6, 9])
def test_witness_satisfies_gates_and_copies(be, k):
    circ = plonk.SyntheticCircuit(k, seed=11 + k)
    cols = [C.fr_rows_to_ints(v) for v in circ.witness(be)]
    u = circ.usable
    a, b, c = cols[0], cols[1], cols[2]
    for i in range(u):
        assert c[i] == (a[i] * b[i] if circ.is_mul[i] else a[i] + b[i]) % O.R_MOD
    flat = np.concatenate([np.array(cols[j][:u] + [0] * (circ.n - u), dtype=object) for j in range(5)])
    # every cell equals the cell it maps to (values are constant along each cycle)
    m = circ.mapping
    for src in range(5 * circ.n):
        assert flat[src] == flat[m[src]]
    # the mapping is a permutation made of the stated cycles
    assert np.array_equal(np.sort(m), np.arange(5 * circ.n))
    assert int((m != np.arange(5 * circ.n)).sum()) == len(circ.copy_cells)
    assert circ.num_cycles == min(1 << 14, 4 * u // 16)


@pytest.mark.parametrize("k", [6, 8])
def test_cpu_proof_verifies(be, k):
    params, circ, pk = setup(be, k, seed=5, tau=0x1234567 + k)
    proof = plonk.create_proof(be, params, pk, circ.witness(be), rng_seed=17)
    assert len(proof) == 13 * 48 + 23 * 32
    assert verify(params, pk, proof)
    # reproducible: the same seeds give the same bytes
    assert plonk.create_proof(be, params, pk, circ.witness(be), rng_seed=17) == proof
    # another rng seed: different blinding, still valid
    other = plonk.create_proof(be, params, pk, circ.witness(be), rng_seed=18)
    assert other != proof and verify(params, pk, other)


def test_rejects_broken_gate_copy_and_tampering(be):
    k = 7
    params, circ, pk = setup(be, k, seed=9, tau=777)
    adv = circ.witness(be)
    good = plonk.create_proof(be, params, pk, adv, rng_seed=3)
    assert verify(params, pk, good)
    bad = [v.copy() for v in adv]
    bad[2][3] = be.add_scalar(bad[2][3:4], 1)[0]
    assert not verify(params, pk, plonk.create_proof(be, params, pk, bad, rng_seed=3))
    cell = int(circ.copy_cells[1])
    slot, row = divmod(cell, circ.usable)
    col = [0, 1, 3, 4][slot]
    bad = [v.copy() for v in adv]
    bad[col][row] = be.add_scalar(bad[col][row:row + 1], 5)[0]
    a_, b_ = bad[0][:circ.usable], bad[1][:circ.usable]
    bad[2][:circ.usable] = be.select(circ.is_mul, be.mul(a_, b_), be.add(a_, b_))
    assert not verify(params, pk, plonk.create_proof(be, params, pk, bad, rng_seed=3))
    for pos in (10, 48 * 5 + 7, len(good) - 100, len(good) - 1):
        t = bytearray(good)
        t[pos] ^= 0x01
        try:
            assert not verify(params, pk, bytes(t))
        except ValueError:
            pass   # not a point / not a canonical scalar: rejected while reading
    assert not verify(params, pk, good + b"\x00")


def test_intermediate_sets_follow_first_appearance():
    qs = [("a", 1, None), ("b", 1, None), ("z", 1, None), ("z", 2, None), ("w", 1, None),
          ("w", 2, None), ("z", 3, None)]
    labels, sets_of, point_sets = plonk.construct_intermediate_sets(qs)
    assert labels == ["a", "b", "z", "w"]
    assert sets_of == [0, 0, 1, 2]
    assert point_sets == [[1], [1, 2, 3], [1, 2]]
    with pytest.raises(ValueError):
        plonk.construct_intermediate_sets([("a", 1, None), ("a", 1, None)])
