"""The benchmark configurations at their full sizes, each stopping at its first
mismatch (GPU):

* 96 batched commitments of floor(2^U) scalars, U ~ U[10, 20], at offsets into one
  resident 2^20 SRS [tau^i]G, each against the closed form [tau^off sum_i s_i tau^i]G
  (kzg/mod.rs:82-108 commit_many over bases[off..]), and the 96 zeta-coset NTTs of
  matching sizes against the C restatement of coeff_to_extended (domain.rs:285-305);
* serving at k in {12, 14, 16, 18, 20}: every served proof equals the same request
  proven alone, and one proof per k, k = 20 included, passes the trapdoor verifier;
* a k = 18 proof on the GPU backend byte-identical to the CPU backend's (the oracle's C
  restatement running the same prover, prover.rs:432-474).
"""
import random
import time

import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests import plonk_verifier as V
from tests.cpu_backend import CpuBackend
from tests.helpers import empty_dev, from_dev, mzk, pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk
serve = pkg.serve


def config4_jobs(n_srs: int, njobs: int = 96, seed: int = 96):
    """Sizes floor(2^U), U ~ U[10, log2 n_srs], and offsets as the benchmark draws them."""
    rng = random.Random(seed)
    hi = n_srs.bit_length() - 1
    sizes = [int(2 ** rng.uniform(10, hi)) for _ in range(njobs)]
    offs = [(i * 7919) % (n_srs - n + 1) for i, n in enumerate(sizes)]
    return sizes, offs


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


@pytest.mark.timeout(600)
def test_config4_batched_commits_closed_form(ctx):
    n_srs = 1 << 20
    sizes, offs = config4_jobs(n_srs)
    total = sum(sizes)
    print(f"\n[config 4: 96 commits, {total} points, resident 2^20 SRS]", flush=True)
    tau = C.fr_from_seed(0xC4, 0, 1)[0]
    tau_i = C.fr_rows_to_ints([tau])[0]
    srs = ctx.srs_generate(tau, n_srs)
    buf = empty_dev(total, 4)
    ctx.fr_from_seed_dev(96, 7, total, buf)
    ctx.synchronize()
    jobs, o = [], 0
    for n, off in zip(sizes, offs):
        jobs.append((buf[o:o + n], n, off))
        o += n
    outs = srs.commit_batch(jobs)
    sc = from_dev(buf).reshape(-1, 4)
    o = 0
    for i, ((_, n, off), out) in enumerate(zip(jobs, outs)):
        s = C.fr_rows_to_ints([C.poly_eval(sc[o:o + n], tau)])[0] * pow(tau_i, off, O.R_MOD) % O.R_MOD
        o += n
        expect = O.g1_affine_from_mont_limbs(C.g1_mul_generator(C.ints_to_fr_rows([s])[0]))
        assert O.g1_jacobian_from_mont_limbs(out) == expect, (i, n, off)
    srs.close()


@pytest.mark.timeout(600)
def test_few_large_commits_against_2p20_srs(ctx):
    """Five 2^17 commits at offsets into a resident 2^20 SRS: shared bucket sets whose
    sort words (24-bit references) fit only with the narrowed second-level key, each
    against the closed form."""
    n_srs, n = 1 << 20, 1 << 17
    tau = C.fr_from_seed(0xC5, 0, 1)[0]
    tau_i = C.fr_rows_to_ints([tau])[0]
    srs = ctx.srs_generate(tau, n_srs)
    buf = empty_dev(5 * n, 4)
    ctx.fr_from_seed_dev(97, 7, 5 * n, buf)
    ctx.synchronize()
    offs = [0, 1, n_srs - n, 333333, 7 * n]
    jobs = [(buf[i * n:(i + 1) * n], n, off) for i, off in enumerate(offs)]
    outs = srs.commit_batch(jobs)
    sc = from_dev(buf).reshape(-1, 4)
    for i, ((_, _, off), out) in enumerate(zip(jobs, outs)):
        s = C.fr_rows_to_ints([C.poly_eval(sc[i * n:(i + 1) * n], tau)])[0] * pow(tau_i, off, O.R_MOD) % O.R_MOD
        expect = O.g1_affine_from_mont_limbs(C.g1_mul_generator(C.ints_to_fr_rows([s])[0]))
        assert O.g1_jacobian_from_mont_limbs(out) == expect, (i, off)
    srs.close()


@pytest.mark.timeout(600)
def test_shared_sets_against_2p23_srs(ctx):
    """Twelve 2^16 commits at offsets into a resident 2^23 SRS (16 GiB of window tables):
    their references need 27 bits, so the low key bits, the reference and its sign no
    longer share one sort word and the scatter stages keys and references apart; each
    commit against the closed form."""
    n_srs, n, nj = 1 << 23, 1 << 16, 12
    print(f"\n[{nj} commits of 2^16 against a resident 2^23 SRS]", flush=True)
    tau = C.fr_from_seed(0xC6, 0, 1)[0]
    tau_i = C.fr_rows_to_ints([tau])[0]
    srs = ctx.srs_generate(tau, n_srs)
    buf = empty_dev(nj * n, 4)
    ctx.fr_from_seed_dev(98, 7, nj * n, buf)
    ctx.synchronize()
    offs = [0, n_srs - n] + [(i * 700001) % (n_srs - n) for i in range(1, nj - 1)]
    jobs = [(buf[i * n:(i + 1) * n], n, off) for i, off in enumerate(offs)]
    outs = srs.commit_batch(jobs)
    assert ctx.msm_last_plan()["bucket_sets"] == nj
    sc = from_dev(buf).reshape(-1, 4)
    for i, ((_, _, off), out) in enumerate(zip(jobs, outs)):
        s = C.fr_rows_to_ints([C.poly_eval(sc[i * n:(i + 1) * n], tau)])[0] * pow(tau_i, off, O.R_MOD) % O.R_MOD
        expect = O.g1_affine_from_mont_limbs(C.g1_mul_generator(C.ints_to_fr_rows([s])[0]))
        assert O.g1_jacobian_from_mont_limbs(out) == expect, (i, off)
    srs.close()


@pytest.mark.timeout(600)
def test_many_shared_sets_commits(ctx):
    """200 commits of 2^15 terms in one batch: 200 shared bucket sets, whose sort bins
    need fewer, larger scalar tiles to keep the (bin, tile) histogram within the scan;
    each against the closed form."""
    n_srs, n, nj = 1 << 16, 1 << 15, 200
    tau = C.fr_from_seed(0xC6, 0, 1)[0]
    tau_i = C.fr_rows_to_ints([tau])[0]
    srs = ctx.srs_generate(tau, n_srs)
    buf = empty_dev(nj * n, 4)
    ctx.fr_from_seed_dev(98, 7, nj * n, buf)
    ctx.synchronize()
    offs = [(i * 331) % (n_srs - n + 1) for i in range(nj)]
    jobs = [(buf[i * n:(i + 1) * n], n, off) for i, off in enumerate(offs)]
    outs = srs.commit_batch(jobs)
    sc = from_dev(buf).reshape(-1, 4)
    for i, ((_, _, off), out) in enumerate(zip(jobs, outs)):
        s = C.fr_rows_to_ints([C.poly_eval(sc[i * n:(i + 1) * n], tau)])[0] * pow(tau_i, off, O.R_MOD) % O.R_MOD
        expect = O.g1_affine_from_mont_limbs(C.g1_mul_generator(C.ints_to_fr_rows([s])[0]))
        assert O.g1_jacobian_from_mont_limbs(out) == expect, (i, off)
    srs.close()


@pytest.mark.timeout(600)
def test_config4_coset_ntts(ctx):
    sizes, _ = config4_jobs(1 << 20)
    print(f"\n[config 4: 96 coset NTTs, {sum(sizes)} coefficients]", flush=True)
    total = sum(sizes)
    buf = empty_dev(total, 4)
    ctx.fr_from_seed_dev(96, 8, total, buf)
    outs, jobs, o = [], [], 0
    for n in sizes:
        log_n = max(1, (n - 1).bit_length())
        d = empty_dev(1 << log_n, 4)
        outs.append(d)
        jobs.append(mzk.NttJob.make(buf[o:o + n], d, log_n, mzk.NTT_COSET, n))
        o += n
    ctx.ntt_batch(jobs)
    ctx.synchronize()
    sc = from_dev(buf).reshape(-1, 4)
    o = 0
    for i, (n, d) in enumerate(zip(sizes, outs)):
        log_n = max(1, (n - 1).bit_length())
        padded = np.zeros((1 << log_n, 4), dtype=np.uint64)
        padded[:n] = sc[o:o + n]
        o += n
        assert np.array_equal(from_dev(d), C.coeff_to_extended(padded, log_n, log_n)), (i, n)


def _verify(params, pk, proof, k):
    fixed = [V.affine_from_limbs(c) for c in pk.fixed_commitments]
    perm = [V.affine_from_limbs(c) for c in pk.permutation_commitments]
    return V.verify(proof, k, pk.transcript_repr, fixed, perm, params.tau)


@pytest.mark.timeout(900)
def test_config5_served_mix_equals_alone_and_verifies():
    ks = (12, 14, 16, 18, 20)
    print("\n[config 5: serving k in 12..20]", flush=True)
    t0 = time.perf_counter()
    server = serve.ProofServer([plonk.GpuBackend() for _ in range(2)], ks=ks, hbm_budget=64 << 30, slots=2)
    print(f"  setup {time.perf_counter() - t0:.1f} s", flush=True)
    reqs = [serve.ProofRequest(rid=i, k=ks[i % 5], witness_seed=900 + i, rng_seed=11 + i) for i in range(10)]
    for engine in ("native", "steps"):
        results, stats = server.run(reqs, realtime=False, workers=2, slots=2, engine=engine)
        assert sorted(r.rid for r in results) == list(range(10)), engine
        assert all(not r.error for r in results), engine
        by = {r.rid: r.proof for r in results}
        if engine == "native":
            native = by
        else:
            assert by == native                       # the two engines agree proof for proof
        print(f"  {engine}: {stats['rounds']} rounds, max in flight {stats['max_inflight']}", flush=True)
    verified = set()
    for r in reqs:
        alone = server.prove_alone(r)
        assert native[r.rid] == alone, (r.rid, r.k)
        if r.k not in verified:
            assert _verify(server.params[r.k], server.pks[r.k], alone, r.k), r.k
            verified.add(r.k)
    assert verified == set(ks)


@pytest.mark.timeout(900)
def test_k18_gpu_proof_identical_to_cpu():
    k = 18
    print("\n[config 3: k=18 GPU vs CPU proof bytes]", flush=True)
    gpu, cpu = plonk.GpuBackend(), CpuBackend(16)
    tau = 0x18C0FFEE
    pg = plonk.Params(gpu, k, tau)
    circ = plonk.SyntheticCircuit(k, seed=18)
    kg = plonk.keygen(gpu, pg, circ)
    wg = circ.witness(gpu)
    proof_g = plonk.create_proof(gpu, pg, kg, wg, rng_seed=99)
    native = plonk.create_proofs_native(gpu, pg, kg, [wg], [99])[0]
    t0 = time.perf_counter()
    pc = plonk.Params(cpu, k, tau)
    kc = plonk.keygen(cpu, pc, circ)
    assert kg.transcript_repr == kc.transcript_repr
    wc = circ.witness(cpu)
    t1 = time.perf_counter()
    proof_c = plonk.create_proof(cpu, pc, kc, wc, rng_seed=99)
    print(f"  cpu setup+keygen {t1 - t0:.1f} s, cpu prove {time.perf_counter() - t1:.1f} s", flush=True)
    assert proof_g == proof_c
    assert native == proof_c
    assert _verify(pg, kg, proof_g, k)
