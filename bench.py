#!/usr/bin/env python3
"""Benchmark of the prover hot path on one MI355X.

A step is one KZG commitment: a G1 MSM of n = 2^20 seeded uniform Fr scalars
against resident SRS bases g_i = [tau^i] G1, with scalars and bases already in
HBM when the timed region starts.  The line printed reports whole-job points/s and
the roofline of the dominant kernel (bucket accumulation) from HIP-event times
measured in this run, and beside it:
  ntt       the scalar-field NTT at 2^22 (HBM and multiply-add fractions);
  batched   96 uneven commitments and 96 coset NTTs, batched against per-call, both in
            the steady state (third call) and with launch graphs off, with roofline
            fractions;
  prove     full k = 18 proofs of the synthetic PLONK circuit (native prover: median ms,
            phase split, lockstep groups; the Python-host prover beside it) and the
            proof's roofline fraction;
  serve     Poisson-arrival serving of k in {12..20}, continuous batching against one
            proof at a time, p50/p99 and roofline fractions;
  cpu_baseline  the oracle's C restatement of the reference CPU path on this host's
            cores (count stated): msm_best at 2^20 (median of 5, which also checks the
            GPU result bit for bit), one thread at 2^16, configuration 1 at 2^16, the
            CPU prover at k = 18 (byte-identical proofs), best_fft at 2^22, the 96-job
            mix (sampled) and sequential CPU proofs/s for the serving mix.
Before anything is timed the autotuner picks the MSM and NTT plans on this device
(--no-tune keeps the models; --tune-file imports a saved table).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--log-n 20] [--no-cpu-baseline] [--cpu-full]
                  [--cpu-prove-k 18] [--no-tune | --tune-file F] [--msm-only | --prove-only | --ntt-only | --serve-only]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "G1 MSM points/s @2^20 + k=18 prove ms; batched vs per-call GPU, bit-exact"
FP_MUL_PER_MADD = 10        # XYZZ mixed addition: 8M + 2S
MAD_PER_FP_MUL = 300        # 12-limb CIOS: 2 * 12^2 + 12 32x32+64 multiply-adds
MAD_PER_FR_MUL = 136        # 8-limb CIOS: 2 * 8^2 + 8
SEED = 0x5EED
COPY_BYTES = 256 << 20      # calibration pass of the traffic probe
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_msm_accum_2p20_traffic.md")


def log(msg: str):
    """Progress of a long run, on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def committed_traffic():
    """HBM bytes per k_msm_accum launch at 2^20 from the committed rocprofv3 counter
    summary (line 'traffic_bytes_per_launch: N'), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def measure_mad_peak(mzk, ctx) -> float:
    """v_mad_u64_u32 issue rate of the chip (mads/s): the INT32 multiply roof."""
    best = 0.0
    for blocks in (256 * 16, 256 * 32):
        mzk.diag_bench(ctx, 4, blocks, 256, 1024)
        ms, ops = mzk.diag_bench(ctx, 4, blocks, 256, 4096)
        best = max(best, ops / (ms * 1e-3))
    return best


def bench_ntt(ctx, torch, peak: float, reps: int = 10) -> dict:
    """2^22 forward NTT on a resident input (the domain's coeff_to_lagrange): wall time per
    transform, the passes' HIP-event time, and their HBM and mad fractions."""
    dom = ctx.domain(1, 22)
    a = torch.zeros((1 << 22, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(SEED, 2, 1 << 22, a)
    for _ in range(2):
        dom.coeff_to_lagrange(a)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        dom.coeff_to_lagrange(a)
    ctx.synchronize()
    ntt_ms = (time.perf_counter() - t1) / reps * 1e3
    ctx.profile(True)
    dom.coeff_to_lagrange(a)
    p2 = ctx.profile_read()
    ctx.profile(False)
    kern_ms = sum(v[0] for k, v in p2.items() if k.startswith("k_ntt_pass"))
    N = 1 << 22
    frmul = (N // 2) * 22
    dom.close()
    return {
        "log_n": 22, "ms": round(ntt_ms, 4), "kernel_ms": round(kern_ms, 4),
        "hbm_GBps_algorithmic": round(64 * N / (kern_ms * 1e-3) / 1e9, 1),
        "hbm_frac_of_8TBps": round(64 * N / (kern_ms * 1e-3) / 8e12, 4),
        "fr_mul_per_s": round(frmul / (kern_ms * 1e-3), 1),
        "mad_frac": round(frmul * MAD_PER_FR_MUL / (kern_ms * 1e-3) / peak, 4),
    }


def msm_mads(n: int) -> int:
    """Algorithmic 32-bit multiply-adds of one MSM of n terms by the reference's window rule:
    c = ceil(ln n), W = floor(255 / c) + 1 windows, W (10 n + 28 2^(c-1)) fp products
    (mixed additions, then summation by parts), 300 mads per fp product."""
    import math
    if n == 0:
        return 0
    c = 1 if n < 4 else (3 if n < 32 else int(math.ceil(math.log(n))))
    w = 255 // c + 1
    return MAD_PER_FP_MUL * w * (10 * n + 28 * (1 << (c - 1)))


def ntt_mads(log_n: int, inverse: bool = False, coset: bool = False) -> int:
    """(N/2) log2 N fr products (+N for the 1/N scaling, +N for the zeta powers), 136 mads each."""
    n = 1 << log_n
    return MAD_PER_FR_MUL * ((n // 2) * log_n + (n if inverse else 0) + (n if coset else 0))


def proof_mads(k: int) -> int:
    """Census of one proof of the synthetic circuit at 2^k rows: 13 commitments of 2^k terms
    (5 advice, 2 permutation products, 4 quotient limbs, f and pi of the multi-open), 7
    inverse NTTs of 2^k, 7 coset NTTs 2^k -> 2^(k+2) and one extended_to_coeff at 2^(k+2)."""
    return (13 * msm_mads(1 << k) + 7 * ntt_mads(k, inverse=True) + 7 * ntt_mads(k + 2, coset=True)
            + ntt_mads(k + 2, inverse=True, coset=True))


def config4_sizes(n_srs: int, njobs: int = 96, seed: int = 96):
    """Sizes floor(2^U), U ~ Uniform[10, log2 n_srs], and the offsets of the jobs."""
    import random
    rng = random.Random(seed)
    hi = n_srs.bit_length() - 1
    sizes = [int(2 ** rng.uniform(10, hi)) for _ in range(njobs)]
    return sizes, [(i * 7919) % (n_srs - n + 1) for i, n in enumerate(sizes)], hi


def bench_batched(mzk, ctx, srs, n_srs: int, peak: float, njobs: int = 96, seed: int = 96) -> dict:
    """96 commitments of sizes floor(2^U), U ~ Uniform[10, log2 n_srs], against the resident
    SRS, plus 96 zeta-coset NTTs of matching sizes (next power of two): per-call GPU
    invocations against one batched invocation, results compared job by job.  Both forms
    are timed in the steady state -- their third call, after a shape's first (eager) call
    and second (graph capture) -- and once more with launch graphs off."""
    import numpy as np
    import torch

    sizes, offs, hi = config4_sizes(n_srs, njobs, seed)
    total = sum(sizes)
    buf = torch.zeros((total, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(seed, 7, total, buf)
    ctx.synchronize()
    views, jobs, o = [], [], 0
    for n, off in zip(sizes, offs):
        v = buf[o:o + n]
        o += n
        views.append(v)
        jobs.append((v, n, off))
    outs = [torch.zeros((1 << max(1, (n - 1).bit_length()), 4), dtype=torch.int64, device="cuda") for n in sizes]
    nj = [mzk.NttJob.make(v, d, max(1, (n - 1).bit_length()), mzk.NTT_COSET, n) for v, d, n in zip(views, outs, sizes)]

    def per_msm():
        return [srs.commit(v, n, offset=off) for v, n, off in jobs]

    def per_ntt():
        for j in nj:
            ctx.ntt(j)

    def timed(fn):
        ctx.synchronize()
        t0 = time.perf_counter()
        r = fn()
        ctx.synchronize()
        return time.perf_counter() - t0, r

    res = {}
    for graphs in (True, False):
        ctx.set_graphs(graphs)
        for _ in range(2):                    # first call eager, second captures its graph
            timed(per_msm)
            timed(lambda: srs.commit_batch(jobs))
            timed(per_ntt)
            timed(lambda: ctx.ntt_batch(nj))
        t_per, per = timed(per_msm)
        t_bat, bat = timed(lambda: srs.commit_batch(jobs))
        t_nper, _ = timed(per_ntt)
        ref = [d.clone() for d in outs]
        for d in outs:
            d.zero_()
        t_nbat, _ = timed(lambda: ctx.ntt_batch(nj))
        if graphs:
            bat_graphs = bat
        res[graphs] = {
            "msm_per_call_ms": round(t_per * 1e3, 3), "msm_batched_ms": round(t_bat * 1e3, 3),
            "msm_speedup": round(t_per / t_bat, 3),
            "ntt_per_call_ms": round(t_nper * 1e3, 3), "ntt_batched_ms": round(t_nbat * 1e3, 3),
            "ntt_speedup": round(t_nper / t_nbat, 3),
            "msm_batched_equals_per_call": all(
                np.array_equal(mzk.g1_to_affine(a), mzk.g1_to_affine(b)) for a, b in zip(per, bat)),
            "ntt_batched_equals_per_call": all(torch.equal(a, b) for a, b in zip(ref, outs)),
        }
    ctx.set_graphs(True)
    g = res[True]
    stats = ctx.graph_stats()
    msm_work = sum(msm_mads(n) for n in sizes)
    ntt_work = sum(ntt_mads(max(1, (n - 1).bit_length()), coset=True) for n in sizes)
    ntt_elems = sum(1 << max(1, (n - 1).bit_length()) for n in sizes)
    return {
        "jobs": njobs, "sizes": "floor(2^U), U~U[10,%d], seed %d" % (hi, seed), "total_points": total,
        "timing": "each form on its third call (steady state: graphs captured where the 48-graph cache holds "
                  "them); 'graphs_off' the same with every launch eager",
        **g,
        "msm_batched_points_per_s": round(total / (g["msm_batched_ms"] * 1e-3), 1),
        "msm_per_call_points_per_s": round(total / (g["msm_per_call_ms"] * 1e-3), 1),
        "ntt_total_elements": ntt_elems,
        "graphs_off": res[False],
        "graph_stats": stats,
        "roofline": {
            "bound": "valu_int32", "unit": "Tmad/s", "peak": round(peak / 1e12, 4),
            "msm_batched": {"mads": msm_work, "achieved": round(msm_work / (g["msm_batched_ms"] * 1e-3) / 1e12, 4),
                            "frac": round(msm_work / (g["msm_batched_ms"] * 1e-3) / peak, 4)},
            "msm_per_call": {"achieved": round(msm_work / (g["msm_per_call_ms"] * 1e-3) / 1e12, 4),
                             "frac": round(msm_work / (g["msm_per_call_ms"] * 1e-3) / peak, 4)},
            "ntt_batched": {"mads": ntt_work, "achieved": round(ntt_work / (g["ntt_batched_ms"] * 1e-3) / 1e12, 4),
                            "frac": round(ntt_work / (g["ntt_batched_ms"] * 1e-3) / peak, 4),
                            "hbm_frac_of_8TBps": round(64 * ntt_elems / (g["ntt_batched_ms"] * 1e-3) / 8e12, 4)},
            "work": "mads by the reference's window rule per job (c = ceil(ln n)) and (N/2) log2 N + N fr "
                    "products per coset NTT; whole-call wall times",
        },
        "_sizes": sizes, "_offs": offs, "_buf": buf, "_bat": bat_graphs,
    }


def bench_prove(pkg, ctx, k: int = 18, reps: int = 5, peak: float = 0.0) -> dict:
    """k-row PLONK/KZG proof of the synthetic circuit on the GPU backend: setup and
    keygen once, witness synthesis, then `reps` proofs (median wall time); the phase
    split comes from one more proof with a device wait at each phase end."""
    import statistics
    plonk = pkg.plonk
    be = plonk.GpuBackend(ctx)
    t0 = time.perf_counter()
    params = plonk.Params(be, k, 0x5EED0000 + k)
    be.sync()
    t1 = time.perf_counter()
    circ = plonk.SyntheticCircuit(k, seed=k)
    pk = plonk.keygen(be, params, circ)
    be.sync()
    t2 = time.perf_counter()
    wit = circ.witness(be)
    be.sync()
    t3 = time.perf_counter()
    proof = plonk.create_proof(be, params, pk, wit, rng_seed=1)
    times = []
    for _ in range(reps):
        be.sync()
        a = time.perf_counter()
        again = plonk.create_proof(be, params, pk, wit, rng_seed=1)
        times.append(time.perf_counter() - a)
        assert again == proof
    phases = {}
    plonk.create_proof(be, params, pk, wit, rng_seed=1, timings=phases)
    prev, split = 0.0, {}
    for name, t in phases.items():
        split[name] = round((t - prev) * 1e3, 3)
        prev = t
    # the library's native prover (C++ host code over the same C ABI): one proof, then a
    # lockstep group of three different witnesses; its proofs must equal the ones above
    native = plonk.create_proofs_native(be, params, pk, [wit], [1])[0]
    ntimes = []
    for _ in range(reps):
        a = time.perf_counter()
        again = plonk.create_proofs_native(be, params, pk, [wit], [1])[0]
        ntimes.append(time.perf_counter() - a)
        assert again == native
    nph = [0.0] * 8
    plonk.create_proofs_native(be, params, pk, [wit], [1], nph)
    # the columns' transforms run on a second stream during the commitments: "transforms_wait"
    # is what remains of them after the permutation commitments ("-" is an empty phase)
    names = ["advice_commit", "permutation", "transforms_wait", "-", "numerator", "quotient_commit", "evaluations",
             "multiopen"]
    nsplit = {nm: round(nph[i] - (nph[i - 1] if i else 0.0), 3) for i, nm in enumerate(names) if nm != "-"}
    wits = [wit] + [circ.witness(be, ws) for ws in (101, 102)]
    seeds = [1, 2, 3]
    group = plonk.create_proofs_native(be, params, pk, wits, seeds)
    gtimes = []
    for _ in range(max(1, reps // 2)):
        a = time.perf_counter()
        plonk.create_proofs_native(be, params, pk, wits, seeds)
        gtimes.append(time.perf_counter() - a)
    group_same = group[0] == proof and all(
        g == plonk.create_proof(be, params, pk, w, rng_seed=sd) for g, w, sd in zip(group[1:], wits[1:], seeds[1:]))
    work = proof_mads(k)
    prove_s = statistics.median(ntimes)
    return {
        "k": k, "prove_ms": round(prove_s * 1e3, 3),
        "roofline": {"bound": "valu_int32", "mads": work, "unit": "Tmad/s",
                     "achieved": round(work / prove_s / 1e12, 4),
                     "frac": round(work / prove_s / peak, 4) if peak else None,
                     "work": "13 commitments of 2^k terms by the reference's window rule (c = ceil(ln n)), "
                             "7 iNTT 2^k, 7 coset NTT 2^k->2^(k+2), 1 extended_to_coeff, per proof; whole-proof "
                             "wall time"},
        "prove_ms_min": round(min(ntimes) * 1e3, 3), "proofs": reps,
        "prover": "native (mzk_plonk_prove)",
        "native_equals_python_host_prover": native == proof,
        "phase_ms": nsplit,
        "group3_ms_per_proof": round(statistics.median(gtimes) * 1e3 / 3, 3),
        "group3_equal_to_single_proofs": group_same,
        "python_host_prover": {"prove_ms": round(statistics.median(times) * 1e3, 3),
                               "prove_ms_min": round(min(times) * 1e3, 3), "phase_ms": split},
        "proof_bytes": len(proof), "msms_per_proof": 13,
        "setup_ms": round((t1 - t0) * 1e3, 1), "keygen_ms": round((t2 - t1) * 1e3, 1),
        "witness_ms": round((t3 - t2) * 1e3, 2),
        "_proof": proof, "_params": params, "_circ": circ,
    }


def bench_serve(pkg, ctx, nreq: int, rates, ks=(12, 14, 16, 18, 20), seed: int = 5,
                workers: int = 4, slots: int = 2, peak: float = 0.0) -> dict:
    """Poisson arrivals of proof requests (k uniform over ks, per-request witness seeds)
    at fixed offered loads, each trace served twice: by the server (`workers` contexts,
    hence streams, with `slots` batched proofs each) and one proof at a time (one
    context, one slot).  A sample of served proofs is compared with the same request
    proven alone."""
    serve = pkg.serve
    plonk = pkg.plonk
    bes = [plonk.GpuBackend(ctx)] + [plonk.GpuBackend() for _ in range(workers - 1)]
    t0 = time.perf_counter()
    server = serve.ProofServer(bes, ks=ks, hbm_budget=64 << 30, slots=slots)
    setup_s = time.perf_counter() - t0
    for be in bes:                      # warm every size on every context (plans, allocator)
        for k in ks:
            server.prove_alone(serve.ProofRequest(rid=-1, k=k, witness_seed=1), be)
        be.sync()
    per_k = {}
    for k in ks:
        t1 = time.perf_counter()
        server.prove_alone(serve.ProofRequest(rid=-1, k=k, witness_seed=2))
        bes[0].sync()
        per_k[k] = round((time.perf_counter() - t1) * 1e3, 2)
    out = {"requests_per_trace": nreq, "ks": list(ks), "workers": workers, "slots_per_worker": slots,
           "alone_ms_by_k": per_k, "setup_s": round(setup_s, 2),
           "hbm_budget": {"configured": server.configured_budget, "measured_free": server.measured_free,
                          "used": server.budget},
           "roofline": {"bound": "valu_int32", "unit": "Tmad/s", "peak": round(peak / 1e12, 4) if peak else None,
                        "work": "per proof the census of the synthetic circuit at its k (13 commitments, 15 NTTs) "
                                "over first arrival to last completion"},
           "loads": []}
    same = True
    for rate in rates:
        trace = serve.poisson_trace(nreq, rate, ks, seed)
        res_n, st_n = server.run(trace, workers=workers, slots=slots, engine="native")
        res_b, st_b = server.run(trace, workers=workers, slots=slots)
        res_s, st_s = server.run(trace, workers=1, slots=1)

        def frac(res):
            # mads of the served proofs over the span from first arrival to last completion
            res = [r for r in res if not r.error]
            st = serve.latency_stats(res)
            span = max(r.done for r in res) - min(r.arrival for r in res)
            rate = sum(proof_mads(r.k) for r in res) / span
            return {**st, "achieved_Tmad_s": round(rate / 1e12, 4),
                    "frac": round(rate / peak, 4) if peak else None}
        out["loads"].append({
            "offered_per_s": rate,
            "served_native": {**frac(res_n), "groups": st_n["rounds"], "max_inflight": st_n["max_inflight"],
                              "rejected": st_n["rejected"]},
            "served": {**frac(res_b), "rounds": st_b["rounds"], "max_inflight": st_b["max_inflight"]},
            "one_at_a_time": frac(res_s),
        })
        byid = {r.rid: r for r in res_s}
        same = same and all(r.proof == byid[r.rid].proof for r in res_b + res_n)
    sample = serve.poisson_trace(nreq, rates[0], ks, seed)[::max(1, nreq // 6)]
    out["served_equals_one_at_a_time"] = same
    out["served_equals_alone_sample"] = all(
        byid_r.proof == server.prove_alone(byid_r_req) for byid_r_req, byid_r in
        ((r, next(x for x in res_b if x.rid == r.rid)) for r in sample))
    return out


def host_cores() -> dict:
    """Host CPUs this process may use: os.cpu_count(), the affinity mask and the cgroup
    CPU quota; `threads` is the smaller of the last two (the CPU legs' pool size)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    use = aff if quota is None else max(1, min(aff, int(quota)))
    return {"nproc": nproc, "affinity": aff, "cgroup_cpu_quota": quota, "threads": use}


def timed_runs(fn, reps: int = 5, warmup: int = 1):
    """Median, min and max seconds of `reps` calls after `warmup` untimed ones, and the
    last result."""
    import statistics
    r = None
    for _ in range(warmup):
        r = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), min(ts), max(ts), r


def cpu_one_thread_leg(C, sc_h, bases_h, cores: int) -> dict:
    """msm_best restatement on one thread at n = 2^16 (bounded: the 2^20 form would take
    about half a minute), median of 3 after a warm-up."""
    n = 1 << 16
    C.set_threads(1)
    try:
        med, mn, mx, _ = timed_runs(lambda: C.msm_best(sc_h[:n], bases_h[:n]), reps=3)
    finally:
        C.set_threads(cores)
    return {"n": n, "cores": 1, "ms": round(med * 1e3, 1), "points_per_s": round(n / med, 1)}


def cpu_config1_2p16(mzk, ctx, srs, sc, C, O) -> dict:
    """Configuration 1 proper: one commitment of 2^16 terms against the first 2^16 SRS
    bases, per-call GPU beside the CPU restatement of msm_best (the reference routes 2^16
    through blst's Pippenger, which cannot be built here), both bit-exact."""
    import numpy as np
    n = 1 << 16
    d = sc[:n]
    ctx.synchronize()
    g_med, g_min, _, out = timed_runs(lambda: srs.commit(d, n), reps=20, warmup=3)
    bases = srs.download(0, n)
    sc_h = d.cpu().numpy().view(np.uint64).copy()
    c_med, c_min, c_max, ref = timed_runs(lambda: C.msm_best(sc_h, bases))
    return {"n": n, "gpu_ms": round(g_med * 1e3, 4), "gpu_points_per_s": round(n / g_med, 1),
            "cpu_ms": round(c_med * 1e3, 2), "cpu_ms_min_max": [round(c_min * 1e3, 2), round(c_max * 1e3, 2)],
            "cpu_points_per_s": round(n / c_med, 1),
            "cpu_routine": "msm_best restatement (the reference uses blst below 2^19 terms; not buildable here)",
            "bit_exact": O.g1_jacobian_from_mont_limbs(ref) == O.g1_jacobian_from_mont_limbs(out)}


def cpu_prove_leg(pkg, k: int, cores: int, gpu_proof: bytes | None, gpu_ms: float | None) -> dict:
    """The same prover on the oracle's C restatement (the CPU prover) at size k, with the
    GPU prover's seeds; the two proofs must be byte-identical."""
    from tests.cpu_backend import CpuBackend
    plonk = pkg.plonk
    cb = CpuBackend(cores)
    t0 = time.perf_counter()
    params = plonk.Params(cb, k, 0x5EED0000 + k)
    circ = plonk.SyntheticCircuit(k, seed=k)
    pk = plonk.keygen(cb, params, circ)
    wit = circ.witness(cb)
    t1 = time.perf_counter()
    proof = plonk.create_proof(cb, params, pk, wit, rng_seed=1)
    cpu_ms = (time.perf_counter() - t1) * 1e3
    out = {"k": k, "cpu_prove_ms": round(cpu_ms, 1), "cpu_setup_keygen_s": round(t1 - t0, 2), "cores": cores}
    if gpu_proof is not None:
        out.update({"gpu_prove_ms": gpu_ms, "identical_proof_bytes": proof == gpu_proof})
    return out


def cpu_ntt_leg(mzk, ctx, C, log_n: int) -> dict:
    """Forward NTT of 2^log_n seeded coefficients by the oracle's C restatement of best_fft
    (bit reverse + DIT), median of 3 after a warm-up, beside the same transform on the GPU;
    the two outputs must be identical."""
    import numpy as np
    import torch
    n = 1 << log_n
    a = torch.zeros((n, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(SEED, 3, n, a)
    out = torch.empty_like(a)
    job = mzk.NttJob.make(a, out, log_n)

    def gpu():
        ctx.ntt(job)
        ctx.synchronize()
    g_med, _, _, _ = timed_runs(gpu, reps=10, warmup=2)
    a_h = a.cpu().numpy().view(np.uint64).copy()
    c_med, c_min, c_max, ref = timed_runs(lambda: C.ntt(a_h, log_n), reps=3)
    same = bool(np.array_equal(ref, out.cpu().numpy().view(np.uint64)))
    return {"log_n": log_n, "cpu_ms": round(c_med * 1e3, 1), "cpu_ms_min_max": [round(c_min * 1e3, 1),
                                                                               round(c_max * 1e3, 1)],
            "gpu_ms": round(g_med * 1e3, 3), "cpu_elements_per_s": round(n / c_med, 1), "bit_exact": same}


def cpu_config4_leg(mzk, C, O, srs, batched: dict, cores: int, full: bool) -> dict:
    """The 96-job mix on the CPU restatement (msm_best per job, the pruned coset DIF per
    transform), each sampled job's result checked against the GPU batched one.  Bounded
    sample: every 6th job (16 of 96) unless full; the total is extrapolated by points."""
    import numpy as np
    sizes, offs = batched["_sizes"], batched["_offs"]
    buf, bat = batched["_buf"], batched["_bat"]
    picks = list(range(len(sizes))) if full else list(range(0, len(sizes), 6))
    starts = np.cumsum([0] + sizes)
    bases = srs.download()
    t_msm = t_ntt = 0.0
    pts = 0
    exact = True
    for i in picks:
        n, off = sizes[i], offs[i]
        sc = buf[starts[i]:starts[i] + n].cpu().numpy().view(np.uint64).copy()
        t0 = time.perf_counter()
        r = C.msm_best(sc, bases[off:off + n])
        t1 = time.perf_counter()
        log_n = max(1, (n - 1).bit_length())
        padded = np.zeros((1 << log_n, 4), dtype=np.uint64)
        padded[:n] = sc
        C.coeff_to_extended(padded, log_n, log_n)
        t2 = time.perf_counter()
        t_msm += t1 - t0
        t_ntt += t2 - t1
        pts += n
        exact = exact and O.g1_jacobian_from_mont_limbs(r) == O.g1_jacobian_from_mont_limbs(bat[i])
    total = sum(sizes)
    scale = total / pts
    return {"jobs_timed": len(picks), "points_timed": pts, "cores": cores,
            "sample": "all 96 jobs" if full else "every 6th job (16 of 96); totals extrapolated by points",
            "msm_ms": round(t_msm * 1e3, 1), "ntt_ms": round(t_ntt * 1e3, 1),
            "msm_total_ms": round(t_msm * scale * 1e3, 1), "ntt_total_ms": round(t_ntt * scale * 1e3, 1),
            "msm_points_per_s": round(pts / t_msm, 1), "sampled_jobs_bit_exact_vs_gpu_batched": exact}


def cpu_config5_leg(pkg, ks, cores: int, known: dict) -> dict:
    """Sequential CPU proving of the serving mix: one CPU proof per k (known[k] reused where
    already measured); proofs/s = 1 / mean over the k drawn uniformly."""
    per_k = {}
    for k in ks:
        per_k[k] = known.get(k) or cpu_prove_leg(pkg, k, cores, None, None)["cpu_prove_ms"]
    mean = sum(per_k.values()) / len(per_k)
    return {"ks": list(ks), "cpu_prove_ms_by_k": per_k, "sequential_proofs_per_s": round(1e3 / mean, 4),
            "cores": cores}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-prove-k", type=int, default=18)
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU legs on every job of the 96-job mix and every served k (k = 20 included): minutes")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[20.0, 40.0, 80.0, 120.0])
    ap.add_argument("--serve-workers", type=int, default=4, help="server contexts (streams)")
    ap.add_argument("--serve-slots", type=int, default=2, help="proofs in flight per server context")
    ap.add_argument("--serve-only", action="store_true", help="only the serving runs (for sweeps of its settings)")
    ap.add_argument("--ntt-only", action="store_true",
                    help="only the NTT figures: 2^22 forward transforms (for rocprofv3 traces of the NTT passes)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run a 256 MiB elementwise copy and 3 MSM commits (for rocprofv3 --pmc runs)")
    ap.add_argument("--prove-only", action="store_true",
                    help="only run keygen and 3 proofs at k=18 (for rocprofv3 kernel traces of the prover)")
    ap.add_argument("--msm-only", action="store_true",
                    help="only the MSM steps and their kernel times (for rocprofv3 kernel traces of the MSM)")
    ap.add_argument("--no-tune", action="store_true", help="keep the planners' built-in models")
    ap.add_argument("--tune-file", default=None,
                    help="import this plan table instead of tuning (the text mzk.tune_export writes)")
    ap.add_argument("--tune-out", default=None, help="write the plan table used to this file")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the MSM/NTT path has no multi-GPU form in this build "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    pkg = importlib.import_module("midnight-zk_amd")
    mzk = pkg.mzk
    ctx = mzk.Context(0)
    # torch work of this process runs on the context's stream: inputs made by torch are
    # ordered before the library's calls with no per-call event
    torch.cuda.set_stream(ctx.torch_stream())
    n = 1 << args.log_n

    if args.prove_only:
        r = bench_prove(pkg, ctx, 18, reps=3, peak=measure_mad_peak(mzk, ctx))
        print(json.dumps({k: v for k, v in r.items() if not k.startswith("_")}), flush=True)
        return 0
    if args.serve_only:
        r = bench_serve(pkg, ctx, args.serve_requests, args.serve_rates, workers=args.serve_workers,
                        slots=args.serve_slots, peak=measure_mad_peak(mzk, ctx))
        print(json.dumps({"serve_only": True, **r}), flush=True)
        return 0

    # ---- resident inputs: tau, SRS bases on the device, scalars on the device
    tau_d = torch.zeros((1, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(SEED, 0, 1, tau_d)
    ctx.synchronize()
    tau = tau_d.cpu().numpy().view(np.uint64)[0].copy()
    t0 = time.perf_counter()
    srs = ctx.srs_generate(tau, n)
    srs_ms = (time.perf_counter() - t0) * 1e3
    sc = torch.zeros((n, 4), dtype=torch.int64, device="cuda")
    ctx.fr_from_seed_dev(SEED, 1, n, sc)
    ctx.synchronize()
    torch.cuda.synchronize()

    # ---- autotuner: MSM plans against this SRS and NTT splits, before anything is timed
    tune = {"mode": "model"}
    if args.tune_file:
        with open(args.tune_file) as f:
            mzk.tune_import(f.read())
        tune = {"mode": "imported", "file": args.tune_file}
    elif not args.no_tune and not args.traffic_probe:
        t0 = time.perf_counter()
        mzk.tune_msm(ctx, srs, 10, args.log_n, reps=3)
        mzk.tune_ntt(ctx, 12, 22, reps=3)
        tune = {"mode": "tuned", "seconds": round(time.perf_counter() - t0, 2)}
    tune["table"] = mzk.tune_export().splitlines()
    if args.tune_out:
        with open(args.tune_out, "w") as f:
            f.write(mzk.tune_export())

    if args.ntt_only:
        # the traffic probe's calibration pass first: rocprofv3 counter runs of this mode
        # scale the passes' FETCH_SIZE / WRITE_SIZE by what the counters read for it
        a = torch.zeros(COPY_BYTES // 8, dtype=torch.int64, device="cuda")
        b = torch.empty_like(a)
        torch.bitwise_xor(a, 1, out=b)
        torch.cuda.synchronize()
        del a, b
        r = bench_ntt(ctx, torch, measure_mad_peak(mzk, ctx), reps=20)
        print(json.dumps({"ntt_only": True, "ntt": r, "tune": tune}), flush=True)
        return 0

    if args.traffic_probe:
        # calibration: one elementwise pass that reads and writes exactly COPY_BYTES each
        a = torch.zeros(COPY_BYTES // 8, dtype=torch.int64, device="cuda")
        b = torch.empty_like(a)
        torch.bitwise_xor(a, 1, out=b)
        torch.cuda.synchronize()
        for _ in range(3):
            srs.commit(sc, n)
        ctx.synchronize()
        print(json.dumps({"traffic_probe": True, "copy_bytes": COPY_BYTES, "msm_n": n}), flush=True)
        return 0

    log("inputs and plans ready; timing the MSM steps")
    # ---- per-call MSM: warmup, then K timed steps between synchronisations
    for _ in range(args.warmup):
        out = srs.commit(sc, n)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = srs.commit(sc, n)
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    plan = ctx.msm_last_plan()

    # ---- per-kernel HIP-event times over 3 more steps (not part of the timed region)
    ctx.profile(True)
    for _ in range(3):
        srs.commit(sc, n)
    prof = ctx.profile_read()
    ctx.profile(False)
    acc_ms = prof["k_msm_accum"][0] / prof["k_msm_accum"][1]
    digits = plan["digits"]
    alg_mads = digits * FP_MUL_PER_MADD * MAD_PER_FP_MUL
    peak = measure_mad_peak(mzk, ctx)
    achieved = alg_mads / (acc_ms * 1e-3)

    if args.msm_only:
        print(json.dumps({"msm_only": True, "n": n, "ms_per_step": round(dt * 1e3, 4),
                          "value": round(n / dt, 1), "plan": plan, "tune": tune,
                          "kernels_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in sorted(prof.items())}}),
              flush=True)
        return 0

    log("MSM steps and kernel times done")
    # ---- NTT at 2^22 beside it (forward, resident input)
    try:
        ntt = bench_ntt(ctx, torch, peak)
    except Exception as e:  # NTT figures are secondary; the MSM line still prints
        ntt = {"error": str(e)}

    log("NTT done")
    # ---- batched commitments and NTTs (many uneven jobs) against per-call invocations
    try:
        batched_raw = bench_batched(mzk, ctx, srs, n, peak)
    except Exception as e:
        batched_raw = {"error": repr(e)}

    log("batched mix done")
    # ---- full proofs at k = 18 on the GPU backend
    try:
        prove_raw = bench_prove(pkg, ctx, 18, peak=peak)
    except Exception as e:
        prove_raw = {"error": repr(e)}

    log("proofs done")
    # ---- serving: Poisson arrivals of mixed-size proofs, continuous batching vs one at a time
    try:
        served = bench_serve(pkg, ctx, args.serve_requests, args.serve_rates, workers=args.serve_workers,
                             slots=args.serve_slots, peak=peak)
    except Exception as e:
        served = {"error": repr(e)}

    log("serving done")
    # ---- CPU baseline: the oracle's C restatement of the reference CPU path on this host's
    # cores, same inputs; every GPU result it sees is checked bit for bit
    cpu = None
    bit_exact = None
    if not args.no_cpu_baseline:
        from oracle import bls12_381 as O
        from oracle import cpu_ref as C
        hc = host_cores()
        cores = args.cpu_threads or hc["threads"]
        C.set_threads(cores)
        bases_h = srs.download()
        sc_h = sc.cpu().numpy().view(np.uint64).copy()
        c_med, c_min, c_max, ref = timed_runs(lambda: C.msm_best(sc_h, bases_h))
        bit_exact = O.g1_jacobian_from_mont_limbs(ref) == O.g1_jacobian_from_mont_limbs(out)
        cpu = {"value": round(n / c_med, 1), "unit": "points/s", "cores": cores, "kind": "port", "host": hc,
               "sample": f"msm_best (C restatement of curves/src/msm.rs:460) at n=2^{args.log_n} on the same bases "
                         f"and scalars: median of 5 runs after 1 warm-up, {c_med:.2f} s "
                         f"(min {c_min:.2f}, max {c_max:.2f})"}
        legs = [
            ("one_thread", lambda: cpu_one_thread_leg(C, sc_h, bases_h, cores)),
            ("config1_2p16", lambda: cpu_config1_2p16(mzk, ctx, srs, sc, C, O)),
            ("prove_sample", lambda: cpu_prove_leg(pkg, args.cpu_prove_k, cores,
                                                   prove_raw.get("_proof") if args.cpu_prove_k == 18 else None,
                                                   prove_raw.get("prove_ms"))),
            ("ntt_sample", lambda: cpu_ntt_leg(mzk, ctx, C, 22)),
            ("config4", lambda: cpu_config4_leg(mzk, C, O, srs, batched_raw, cores, args.cpu_full)),
        ]
        log("CPU msm_best done")
        for name, fn in legs:
            try:
                cpu[name] = fn()
                log(f"CPU leg {name} done")
            except Exception as e:
                cpu[name] = {"error": repr(e)}
        try:
            known = {}
            ps = cpu.get("prove_sample", {})
            if "cpu_prove_ms" in ps:
                known[ps["k"]] = ps["cpu_prove_ms"]
            ks = (12, 14, 16, 18, 20) if args.cpu_full else (12, 14, 16, 18)
            c5 = cpu_config5_leg(pkg, ks, cores, known)
            if not args.cpu_full:
                c5["sample"] = "k in {12, 14, 16, 18} (k = 20 omitted to bound the run: it would lower the rate)"
            if isinstance(served, dict) and served.get("loads"):
                best = max(ld["served_native"]["proofs_per_s"] for ld in served["loads"])
                c5["gpu_served_native_best_proofs_per_s"] = best
            cpu["config5"] = c5
        except Exception as e:
            cpu["config5"] = {"error": repr(e)}

    batched = {k: v for k, v in batched_raw.items() if not k.startswith("_")}
    prove = {k: v for k, v in prove_raw.items() if not k.startswith("_")}

    line = {
        "metric": METRIC,
        "value": round(n / dt, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"G1 MSM n=2^{args.log_n} (one KZG commit) against resident SRS bases [tau^i]G1, "
                        f"uniform seeded Fr scalars, inputs resident in HBM",
            "n": n, "window_bits": plan["c"], "windows": plan["windows"],
            "bucket_sets": plan["bucket_sets"], "refs_per_thread": plan["refs_per_thread"],
            "parallelism": "1 GPU",
        },
        "roofline": {
            "kernel": "k_msm_accum",
            "bound": "valu_int32",
            "achieved": round(achieved / 1e12, 4),
            "peak": round(peak / 1e12, 4),
            "unit": "Tmad/s",
            "frac": round(achieved / peak, 4),
            "traffic": committed_traffic() if args.log_n == 20 else None,
            "per_launch": {"madds": digits, "mads_algorithmic": alg_mads, "avg_ms": round(acc_ms, 4)},
        },
        "kernels_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in sorted(prof.items())},
        "cpu_baseline": cpu,
        "bit_exact_vs_cpu": bit_exact,
        "ntt": ntt,
        "batched": batched,
        "prove": prove,
        "serve": served,
        "setup": {"srs_generate_ms": round(srs_ms, 2)},
        "tune": tune,
    }
    print(json.dumps(line), flush=True)
    # release everything built on the context (proof parameters, batch buffers) before it
    del batched_raw, prove_raw
    srs.close()
    ctx.close()
    return 0 if bit_exact is not False else 1


if __name__ == "__main__":
    sys.exit(main())
