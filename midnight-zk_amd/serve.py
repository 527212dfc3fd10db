"""Proof serving with continuous batching on one GPU.

Requests (k, witness seed, arrival time) arrive over time.  The server keeps one set
of monomial SRS bases (2^max_k points) and, per supported k, the Lagrange bases, the
evaluation domain and the proving key resident on the device.  A request is admitted
as soon as it has arrived, a worker has a free slot, and its estimated device footprint
fits in the HBM budget beside the proofs already in flight (first come, first served).
Every in-flight proof is a prove_steps generator; each scheduling round of a worker
collects the pending request of each of its in-flight proofs and answers all
commitments with one batched MSM pipeline per SRS, all transforms with one batched NTT
launch pair, and all evaluations with one launch.  Newly arrived requests join at the
next round.  A proof's bytes do not depend on what else is in flight: every
batched operation returns exactly what the per-call operation returns.

Several workers, each with its own library context (its own HIP stream), serve at
once: their kernels run concurrently on the device, and one worker's host-side work and
waits overlap the others' device work.

The reference has no serving layer; this is the request loop around its prover
(create_proof, proofs/src/plonk/prover.rs:432) for many concurrent callers.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

from . import plonk


@dataclass
class ProofRequest:
    rid: int
    k: int
    witness_seed: int
    arrival: float = 0.0          # seconds after the start of the run
    rng_seed: int = 1


@dataclass
class ProofResult:
    rid: int
    k: int
    arrival: float
    admitted: float
    done: float
    proof: bytes = field(repr=False, default=b"")

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def footprint_bytes(k: int) -> int:
    """Device bytes one in-flight proof of size 2^k holds at its peak: about 24 vectors of
    2^k scalars and 14 of 2^(k+2) (advice and permutation columns, their coefficient and
    extended forms, the numerator and quotient), 32 bytes each."""
    n = 1 << k
    return (24 * n + 14 * 4 * n) * 32


class ProofServer:
    """backends: one prover backend per worker.  Each GPU worker owns a library context,
    hence its own HIP stream and scratch memory; the resident SRS bases and proving keys
    are built once with backends[0] and shared read-only.  Each worker drives up to
    `slots` in-flight proofs in batched rounds; the workers run in their own threads,
    so their device work overlaps on the GPU and their host waits overlap each other."""

    def __init__(self, backends, ks=(12, 14, 16, 18, 20), tau: int = 0x5EEDC0DE, layout_seed: int = 1,
                 hbm_budget: int = 32 << 30, slots: int = 2):
        if not isinstance(backends, (list, tuple)):
            backends = [backends]
        self.backends = list(backends)
        be = self.be = self.backends[0]
        self.ks = sorted(set(ks))
        self.budget = hbm_budget
        self.slots = slots
        kmax = self.ks[-1]
        self.g = be.srs_monomial(tau % plonk.R_MOD, 1 << kmax)
        self.params, self.pks, self.circuits = {}, {}, {}
        for k in self.ks:
            p = plonk.Params(be, k, tau, g=self.g)
            c = plonk.SyntheticCircuit(k, layout_seed + k)
            self.params[k], self.circuits[k] = p, c
            self.pks[k] = plonk.keygen(be, p, c)
        be.sync()

    def prove_alone(self, req: ProofRequest, be=None) -> bytes:
        be = be or self.be
        k = req.k
        wit = self.circuits[k].witness(be, req.witness_seed)
        return plonk.create_proof(be, self.params[k], self.pks[k], wit, req.rng_seed)

    def run(self, requests, realtime: bool = True, workers: int | None = None, slots: int | None = None):
        """Serves the requests (arrival times relative to the start of the run; with
        realtime=False every request counts as arrived at once) on the first `workers`
        backends with `slots` proofs per worker.  Returns ProofResults in completion
        order and the run's statistics."""
        import threading
        for r in requests:
            if r.k not in self.pks:
                raise ValueError(f"k={r.k} is not served (served: {self.ks})")
        bes = self.backends[:workers or len(self.backends)]
        slots = slots or self.slots
        queue = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        lock = threading.Lock()
        state = {"used": 0, "inflight": 0, "max_inflight": 0, "rounds": 0}
        done, errors = [], []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))

        def admit(n_local):
            """Pops the requests this worker may start now (first come, first served)."""
            out = []
            with lock:
                now = clock()
                while (queue and queue[0].arrival <= now and n_local + len(out) < slots
                       and (state["inflight"] == 0 or state["used"] + footprint_bytes(queue[0].k) <= self.budget)):
                    req = queue.popleft()
                    state["used"] += footprint_bytes(req.k)
                    state["inflight"] += 1
                    state["max_inflight"] = max(state["max_inflight"], state["inflight"])
                    out.append(req)
                nxt = queue[0].arrival if queue else None
            return out, nxt

        def worker(be):
            inflight = []
            try:
                while True:
                    new, nxt = admit(len(inflight))
                    for req in new:
                        k = req.k
                        wit = self.circuits[k].witness(be, req.witness_seed)
                        gen = plonk.prove_steps(be, self.params[k], self.pks[k], wit, req.rng_seed)
                        inflight.append([req, gen, gen.send(None), time.perf_counter() - t0])
                    if not inflight:
                        if nxt is None:
                            with lock:
                                if not queue:
                                    return
                            continue
                        time.sleep(max(0.0002, min(nxt - clock(), 0.002)))
                        continue
                    with lock:
                        state["rounds"] += 1
                    by_kind = collections.defaultdict(list)
                    for slot in inflight:
                        kind, items = slot[2]
                        by_kind[kind].append((slot, items))
                    answers = {}
                    for kind, parts in by_kind.items():
                        res = plonk.serve_request(be, (kind, [it for _, items in parts for it in items]))
                        o = 0
                        for slot, items in parts:
                            answers[id(slot)] = res[o:o + len(items)]
                            o += len(items)
                    still = []
                    for slot in inflight:
                        req, gen = slot[0], slot[1]
                        try:
                            slot[2] = gen.send(answers[id(slot)])
                            still.append(slot)
                        except StopIteration as stop:
                            be.sync()
                            t = time.perf_counter() - t0
                            with lock:
                                done.append(ProofResult(req.rid, req.k, req.arrival, slot[3], t, stop.value))
                                state["used"] -= footprint_bytes(req.k)
                                state["inflight"] -= 1
                    inflight = still
            except BaseException as e:  # surfaced to the caller after the join
                errors.append(e)
                with lock:
                    queue.clear()

        threads = [threading.Thread(target=worker, args=(be,), daemon=True) for be in bes]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        wall = time.perf_counter() - t0
        return done, {"rounds": state["rounds"], "max_inflight": state["max_inflight"], "wall_s": wall,
                      "workers": len(bes), "slots": slots}


def poisson_trace(n: int, rate: float, ks, seed: int):
    """n requests with exponential inter-arrival times of mean 1/rate seconds and k drawn
    uniformly from ks (seeded)."""
    import random
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.expovariate(rate)
        out.append(ProofRequest(rid=i, k=rng.choice(list(ks)), witness_seed=seed * 100003 + i, arrival=t,
                                rng_seed=i + 1))
    return out


def latency_stats(results) -> dict:
    lat = sorted(r.latency for r in results)
    if not lat:
        return {}

    def pct(p):
        return lat[min(len(lat) - 1, int(round(p / 100.0 * (len(lat) - 1))))]
    first = min(r.arrival for r in results)
    last = max(r.done for r in results)
    return {"proofs": len(lat), "proofs_per_s": round(len(lat) / (last - first), 3),
            "p50_ms": round(pct(50) * 1e3, 2), "p99_ms": round(pct(99) * 1e3, 2),
            "mean_ms": round(sum(lat) / len(lat) * 1e3, 2), "max_ms": round(lat[-1] * 1e3, 2)}
