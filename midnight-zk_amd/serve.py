"""Proof serving with continuous batching on one GPU.

Requests (k, witness seed, arrival time) arrive over time.  The server keeps one set
of monomial SRS bases (2^max_k points) and, per supported k, the Lagrange bases, the
evaluation domain and the proving key resident on the device.  A request is admitted
as soon as it has arrived, a worker has a free slot, and its estimated device footprint
fits in the HBM budget beside the proofs already in flight (first come, first served).
The budget is the smaller of the configured one and the device memory measured free
(hipMemGetInfo) once the resident SRS and keys are in place, less a reserve; a request
whose footprint alone exceeds the budget is rejected at once with a status, never
started.
Every in-flight proof is a prove_steps generator; each scheduling round of a worker
collects the pending request of each of its in-flight proofs and answers all
commitments with one batched MSM pipeline per SRS, all transforms with one batched NTT
launch pair, and all evaluations with one launch.  Newly arrived requests join at the
next round.  A proof's bytes do not depend on what else is in flight: every
batched operation returns exactly what the per-call operation returns.

Several workers, each with its own library context (its own HIP stream), serve at
once: their kernels run concurrently on the device, and one worker's host-side work and
waits overlap the others' device work.  With the native engine a worker proves the
oldest arrived request together with the arrived requests of the same size as one
lockstep group in the library's C++ prover (mzk_plonk_prove), outside the interpreter
lock.

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
    error: str = ""               # non-empty: the request was rejected (and not proven)

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

    # device memory left unclaimed by admission: scratch growth, plans and graphs
    RESERVE_BYTES = 4 << 30

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
        self.configured_budget = hbm_budget
        self.measured_free = None
        mem_info = getattr(be, "mem_info", None)
        if mem_info is not None:
            self.measured_free = mem_info()[0]
            self.budget = min(self.budget, max(0, self.measured_free - self.RESERVE_BYTES))

    def prove_alone(self, req: ProofRequest, be=None) -> bytes:
        be = be or self.be
        k = req.k
        wit = self.circuits[k].witness(be, req.witness_seed)
        return plonk.create_proof(be, self.params[k], self.pks[k], wit, req.rng_seed)

    def run(self, requests, realtime: bool = True, workers: int | None = None, slots: int | None = None,
            engine: str = "steps"):
        """Serves the requests (arrival times relative to the start of the run; with
        realtime=False every request counts as arrived at once) on the first `workers`
        backends with `slots` proofs per worker.  engine "steps": each worker drives its
        in-flight proofs as prove_steps generators in batched rounds (requests join at
        the next round); engine "native": each worker takes the oldest arrived request
        together with the arrived requests of the same size (up to `slots`) and proves
        them as one lockstep group in the library's native prover, which runs without
        the interpreter lock, so the workers' host work overlaps too.  Returns
        ProofResults in completion order and the run's statistics."""
        if engine not in ("steps", "native"):
            raise ValueError(f"unknown engine {engine!r}")
        import threading
        for r in requests:
            if r.k not in self.pks:
                raise ValueError(f"k={r.k} is not served (served: {self.ks})")
        bes = self.backends[:workers or len(self.backends)]
        slots = slots or self.slots
        queue = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        lock = threading.Lock()
        state = {"used": 0, "inflight": 0, "max_inflight": 0, "rounds": 0, "rejected": 0}
        done, errors = [], []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))

        def reject_oversize(now):
            """Requests at the head of the queue that no budget state could ever admit are
            answered at once with an error (caller holds the lock)."""
            while queue and queue[0].arrival <= now and footprint_bytes(queue[0].k) > self.budget:
                req = queue.popleft()
                t = time.perf_counter() - t0
                done.append(ProofResult(req.rid, req.k, req.arrival, t, t, b"",
                                        f"rejected: footprint {footprint_bytes(req.k)} B exceeds the HBM budget "
                                        f"{self.budget} B"))
                state["rejected"] += 1

        def admit(n_local):
            """Pops the requests this worker may start now (first come, first served)."""
            out = []
            with lock:
                now = clock()
                reject_oversize(now)
                while (queue and queue[0].arrival <= now and n_local + len(out) < slots
                       and state["used"] + footprint_bytes(queue[0].k) <= self.budget):
                    req = queue.popleft()
                    state["used"] += footprint_bytes(req.k)
                    state["inflight"] += 1
                    state["max_inflight"] = max(state["max_inflight"], state["inflight"])
                    out.append(req)
                    reject_oversize(now)
                nxt = queue[0].arrival if queue else None
            return out, nxt

        def admit_group():
            """Native engine: the oldest arrived request and the arrived ones of its size."""
            with lock:
                now = clock()
                reject_oversize(now)
                if not queue or queue[0].arrival > now:
                    return [], (queue[0].arrival if queue else None)
                k = queue[0].k
                picked = []
                for r in queue:
                    if r.arrival > now or len(picked) == slots:
                        break
                    if r.k != k:
                        continue
                    if state["used"] + footprint_bytes(k) * (len(picked) + 1) > self.budget:
                        break
                    picked.append(r)
                if not picked:
                    return [], (queue[0].arrival if queue else None)
                for r in picked:
                    queue.remove(r)
                state["used"] += footprint_bytes(k) * len(picked)
                state["inflight"] += len(picked)
                state["max_inflight"] = max(state["max_inflight"], state["inflight"])
                state["rounds"] += 1
                return picked, None

        def native_worker(be):
            try:
                while True:
                    group, nxt = admit_group()
                    if not group:
                        with lock:
                            if not queue:
                                return
                        wait = 0.0005 if nxt is None else max(0.0002, min(nxt - clock(), 0.002))
                        time.sleep(wait)
                        continue
                    start = time.perf_counter() - t0
                    k = group[0].k
                    wits = [self.circuits[k].witness(be, r.witness_seed) for r in group]
                    proofs = plonk.create_proofs_native(be, self.params[k], self.pks[k], wits,
                                                        [r.rng_seed for r in group])
                    t = time.perf_counter() - t0
                    with lock:
                        for r, pr in zip(group, proofs):
                            done.append(ProofResult(r.rid, r.k, r.arrival, start, t, pr))
                        state["used"] -= footprint_bytes(k) * len(group)
                        state["inflight"] -= len(group)
            except BaseException as e:  # surfaced to the caller after the join
                errors.append(e)
                with lock:
                    queue.clear()

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

        target = native_worker if engine == "native" else worker
        threads = [threading.Thread(target=target, args=(be,), daemon=True) for be in bes]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        wall = time.perf_counter() - t0
        return done, {"rounds": state["rounds"], "max_inflight": state["max_inflight"], "wall_s": wall,
                      "workers": len(bes), "slots": slots, "engine": engine, "rejected": state["rejected"],
                      "budget": self.budget}


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
    results = [r for r in results if not r.error]
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
