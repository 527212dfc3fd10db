"""Launch sequences replayed from cached HIP graphs give the eager results: MSMs on both
bucket layouts and NTT batches, each replayed several times with other eager work and
waits between the replays, and native proofs with graphs on and off (GPU)."""
import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import empty_dev, from_dev, mzk, pkg, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def aff(j):
    return O.g1_jacobian_from_mont_limbs(j)


def test_msm_and_ntt_graph_replays_equal_eager(ctx):
    tau = C.fr_from_seed(5, 0, 1)[0]
    srs = ctx.srs_upload(C.srs_bases(tau, 1 << 16))
    big = to_dev(C.fr_from_seed(31, 1, 1 << 16))                 # shared bucket set
    small = [to_dev(C.fr_from_seed(40 + i, 1, 3000)) for i in range(5)]   # per-window small sets
    small_jobs = [(d, 3000, 100 * i) for i, d in enumerate(small)]
    specs = [(12, mzk.NTT_COSET, 1 << 10), (14, mzk.NTT_INVERSE, 1 << 14), (16, 0, 1 << 16)]
    ins = [to_dev(C.fr_from_seed(50 + i, lg, nr)) for i, (lg, _, nr) in enumerate(specs)]
    outs = [empty_dev(1 << lg, 4) for lg, _, _ in specs]
    jobs = [mzk.NttJob.make(a, o, lg, fl, nr) for a, o, (lg, fl, nr) in zip(ins, outs, specs)]
    other = empty_dev(1 << 12, 4)

    ctx.set_graphs(False)
    want_big = aff(srs.commit(big, 1 << 16))
    want_small = [aff(p) for p in srs.commit_batch(small_jobs)]
    ctx.ntt_batch(jobs)
    ctx.synchronize()
    want_ntt = [from_dev(o).copy() for o in outs]

    ctx.set_graphs(True)
    for rep in range(5):   # a shape runs eagerly once, is captured on its second call, then replays
        print(f"\n[graphs] replay {rep}", flush=True)
        for o in outs:
            o.zero_()
        assert aff(srs.commit(big, 1 << 16)) == want_big
        ctx.fr_from_seed_dev(900 + rep, 2, 1 << 12, other)          # eager work between replays
        ctx.synchronize()
        assert [aff(p) for p in srs.commit_batch(small_jobs)] == want_small
        ctx.ntt_batch(jobs)
        ctx.synchronize()
        for o, w in zip(outs, want_ntt):
            assert np.array_equal(from_dev(o), w)
    st = ctx.graph_stats()
    assert st["captured"] >= 3 and st["replayed"] >= 9, st
    srs.close()


def test_native_proofs_with_and_without_graphs(ctx):
    plonk = pkg.plonk
    be = plonk.GpuBackend(ctx)
    k = 8
    params = plonk.Params(be, k, 0x6A6A + k)
    circ = plonk.SyntheticCircuit(k, seed=3)
    pk = plonk.keygen(be, params, circ)
    wits = [circ.witness(be, 70 + i) for i in range(2)]
    ctx.set_graphs(False)
    eager = plonk.create_proofs_native(be, params, pk, wits, [4, 5])
    ctx.set_graphs(True)
    for rep in range(3):
        print(f"\n[graphs] native group {rep}", flush=True)
        assert plonk.create_proofs_native(be, params, pk, wits, [4, 5]) == eager
    assert ctx.graph_stats()["replayed"] > 0


def test_working_set_above_cache_size(ctx):
    # 60 shapes cycled four times: more shapes than the 48-graph cache holds.  Results stay
    # the eager ones, the cache keeps its first 48 shapes (no recapture on every call) and
    # the shapes beyond them run eagerly.
    tau = C.fr_from_seed(8, 0, 1)[0]
    srs = ctx.srs_upload(C.srs_bases(tau, 1 << 15))
    sizes = [1000 + 37 * i for i in range(60)]
    sc = [to_dev(C.fr_from_seed(600 + i, 1, n)) for i, n in enumerate(sizes)]
    ctx.set_graphs(False)
    want = [aff(srs.commit(d, n)) for d, n in zip(sc, sizes)]
    ctx.set_graphs(True)          # frees every graph: counters below start from this state
    base = ctx.graph_stats()
    for rep in range(4):
        print(f"\n[graphs] 60-shape cycle {rep}", flush=True)
        for d, n, w in zip(sc, sizes, want):
            assert aff(srs.commit(d, n)) == w, (rep, n)
    st = ctx.graph_stats()
    captured = st["captured"] - base["captured"]
    replayed = st["replayed"] - base["replayed"]
    assert st["held"] <= 48 and captured <= 48, st
    assert replayed >= 2 * 48, st        # cycles 3 and 4 replay the cached shapes
    srs.close()
