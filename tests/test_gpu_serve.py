"""The proof server on the GPU: proofs served in a mix of sizes, with their MSMs, NTTs
and evaluations batched across in-flight proofs, are byte-identical to the same
requests proven alone and pass the trapdoor verifier (GPU)."""
import pytest

from tests import plonk_verifier as V
from tests.helpers import pkg

pytestmark = pytest.mark.gpu
serve = pkg.serve
plonk = pkg.plonk


@pytest.fixture(scope="module")
def server():
    return serve.ProofServer([plonk.GpuBackend() for _ in range(3)], ks=(8, 10, 12), hbm_budget=8 << 30, slots=4)


def test_mix_equals_alone_and_verifies(server):
    print("\n[serve mix k=8,10,12]", flush=True)
    reqs = [serve.ProofRequest(rid=i, k=(8, 10, 12)[i % 3], witness_seed=500 + i, rng_seed=3 + i) for i in range(9)]
    results, stats = server.run(reqs, realtime=False, workers=1, slots=16)
    assert stats["max_inflight"] == 9
    for r in results:
        req = reqs[r.rid]
        assert r.proof == server.prove_alone(req), r.rid
        pk, params = server.pks[req.k], server.params[req.k]
        fixed = [V.affine_from_limbs(c) for c in pk.fixed_commitments]
        perm = [V.affine_from_limbs(c) for c in pk.permutation_commitments]
        assert V.verify(r.proof, req.k, pk.transcript_repr, fixed, perm, params.tau)


def test_poisson_arrivals(server):
    print("\n[serve poisson 24 requests]", flush=True)
    reqs = serve.poisson_trace(24, 60.0, (8, 10, 12), seed=9)
    results, stats = server.run(reqs, workers=3, slots=2)
    assert len(results) == 24
    st = serve.latency_stats(results)
    print(f"  {st} rounds={stats['rounds']} max_inflight={stats['max_inflight']}", flush=True)
    for r in results[:6]:
        assert r.proof == server.prove_alone(reqs[r.rid])


def test_native_engine_groups_equal_alone(server):
    print("\n[serve native engine, all arrived]", flush=True)
    reqs = [serve.ProofRequest(rid=i, k=(8, 10, 12)[i % 3], witness_seed=700 + i, rng_seed=5 + i) for i in range(12)]
    results, stats = server.run(reqs, realtime=False, workers=2, slots=3, engine="native")
    assert len(results) == 12 and stats["engine"] == "native"
    assert stats["rounds"] <= 6            # same-size requests share lockstep groups
    for r in results:
        assert r.proof == server.prove_alone(reqs[r.rid]), r.rid


def test_native_engine_poisson(server):
    print("\n[serve native engine, poisson 24 requests]", flush=True)
    reqs = serve.poisson_trace(24, 80.0, (8, 10, 12), seed=13)
    results, stats = server.run(reqs, workers=3, slots=2, engine="native")
    assert sorted(r.rid for r in results) == list(range(24))
    print(f"  {serve.latency_stats(results)} groups={stats['rounds']}", flush=True)
    for r in results[::4]:
        assert r.proof == server.prove_alone(reqs[r.rid])
