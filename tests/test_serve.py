"""The proof server's scheduling on the CPU backend: every proof served in a mix of
sizes is byte-identical to the same request proven alone, admission respects the
memory budget, and the Poisson trace is reproducible."""
import pytest

from tests.cpu_backend import CpuBackend
from tests.helpers import pkg

serve = pkg.serve


@pytest.fixture(scope="module")
def server():
    return serve.ProofServer([CpuBackend(8), CpuBackend(8)], ks=(5, 6, 7), hbm_budget=1 << 30, slots=8)


def test_mix_equals_alone(server):
    reqs = [serve.ProofRequest(rid=i, k=(5, 6, 7)[i % 3], witness_seed=100 + i, rng_seed=7 + i)
            for i in range(7)]
    results, stats = server.run(reqs, realtime=False, workers=1)
    assert sorted(r.rid for r in results) == list(range(7))
    assert stats["max_inflight"] == 7          # all admitted together
    for r in results:
        assert r.proof == server.prove_alone(reqs[r.rid]), r.rid
    results, stats = server.run(reqs, realtime=False, workers=2, slots=2)
    assert sorted(r.rid for r in results) == list(range(7)) and stats["max_inflight"] <= 4
    for r in results:
        assert r.proof == server.prove_alone(reqs[r.rid]), r.rid


def test_budget_limits_inflight(server):
    fp = serve.footprint_bytes(7)
    server.budget = 2 * fp
    try:
        reqs = [serve.ProofRequest(rid=i, k=7, witness_seed=i) for i in range(5)]
        results, stats = server.run(reqs, realtime=False, workers=2, slots=8)
        assert stats["max_inflight"] == 2 and len(results) == 5
    finally:
        server.budget = 1 << 30


def test_oversize_request_rejected_even_when_idle(server):
    # a request whose footprint alone exceeds the budget is answered with an error at once
    # and never started; the others are served as usual
    server.budget = serve.footprint_bytes(6)
    try:
        reqs = [serve.ProofRequest(rid=0, k=5, witness_seed=1), serve.ProofRequest(rid=1, k=7, witness_seed=2),
                serve.ProofRequest(rid=2, k=6, witness_seed=3)]
        results, stats = server.run(reqs, realtime=False, workers=1, slots=4)
        by = {r.rid: r for r in results}
        assert sorted(by) == [0, 1, 2] and stats["rejected"] == 1
        assert by[1].error.startswith("rejected") and by[1].proof == b""
        for rid in (0, 2):
            assert not by[rid].error and by[rid].proof == server.prove_alone(reqs[rid])
        assert serve.latency_stats(results)["proofs"] == 2
    finally:
        server.budget = 1 << 30


def test_unknown_size_rejected(server):
    with pytest.raises(ValueError):
        server.run([serve.ProofRequest(rid=0, k=9, witness_seed=1)], realtime=False)


def test_poisson_trace_and_stats():
    a = serve.poisson_trace(50, 20.0, (12, 14), seed=3)
    b = serve.poisson_trace(50, 20.0, (12, 14), seed=3)
    assert [(r.k, r.arrival) for r in a] == [(r.k, r.arrival) for r in b]
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    res = [serve.ProofResult(i, 12, float(i), float(i), i + 0.01 * (i + 1)) for i in range(100)]
    st = serve.latency_stats(res)
    assert st["p50_ms"] == pytest.approx(10 * 51, rel=0.03) and st["p99_ms"] == pytest.approx(990, rel=0.02)
