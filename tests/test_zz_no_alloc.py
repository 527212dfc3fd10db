"""mzk_ntt_dev_noalloc, the NTT entry point that allocates nothing, against mzk_ntt_dev
(the allocating entry point it is a variant of) through the C ABI: bit-identical outputs
on two calls in a row, the second with a smaller input, and the size limit (GPU)."""
import numpy as np
import pytest

from oracle import cpu_ref as C
from tests.helpers import empty_dev, from_dev, mzk, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def _both(ctx, log_n, flags, n_real, seed):
    vals = C.fr_from_seed(seed, log_n, n_real)
    din = to_dev(vals)
    ref = empty_dev(1 << log_n, 4)
    got = empty_dev(1 << log_n, 4)
    ctx.ntt(mzk.NttJob.make(din, ref, log_n, flags, n_real))
    ctx.ntt_noalloc(mzk.NttJob.make(din, got, log_n, flags, n_real))
    ctx.synchronize()
    return from_dev(ref), from_dev(got)


def test_noalloc_equals_allocating_entry_point(ctx):
    # the largest size first, then a smaller one on the same (already used) device array,
    # then the single-pass size range
    for log_n, flags, n_real, seed in [(20, mzk.NTT_COSET, (1 << 20) - 5, 1),
                                       (16, mzk.NTT_INVERSE | mzk.NTT_COSET, 1 << 16, 2),
                                       (9, 0, 300, 3)]:
        print(f"\n[noalloc 2^{log_n} flags {flags}]", flush=True)
        ref, got = _both(ctx, log_n, flags, n_real, seed)
        assert np.array_equal(got, ref), (log_n, flags)
        assert ref.any()


def test_noalloc_in_place_and_limit(ctx):
    vals = C.fr_from_seed(4, 4, 1 << 18)
    d = to_dev(vals)
    ctx.ntt_noalloc(mzk.NttJob.make(d, d, 18))
    ctx.synchronize()
    assert np.array_equal(from_dev(d), C.ntt(vals, 18))
    big = empty_dev(16, 4)
    with pytest.raises(mzk.MzkError, match="maximum"):
        ctx.ntt_noalloc(mzk.NttJob.make(big, big, mzk.NTT_NOALLOC_MAX_LOG + 1, 0, 16))
