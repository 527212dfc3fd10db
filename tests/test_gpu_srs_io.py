"""SRS files on the device: ParamsKZG::write_custom / read_custom's G1 point arrays
(ZCash compressed and uncompressed encodings) encoded and decoded by the HIP kernels,
byte for byte against the oracle's encoders, point for point against the oracle's
decoders, every rejection class with the lowest failing index reported, and a whole
parameters file read, its delta bases derived, and written back unchanged (GPU)."""
import numpy as np
import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C
from tests.helpers import mzk, pkg

pytestmark = pytest.mark.gpu
plonk = pkg.plonk
FORMATS = [mzk.SERDE_PROCESSED, mzk.SERDE_RAW_BYTES, mzk.SERDE_RAW_BYTES_UNCHECKED]


@pytest.fixture(scope="module")
def ctx():
    C.set_threads(16)
    c = mzk.Context(0)
    yield c
    c.close()


def encode(P, fmt):
    return O.g1_to_compressed(P) if fmt == mzk.SERDE_PROCESSED else O.g1_to_uncompressed(P)


def decode(b, fmt):
    if fmt == mzk.SERDE_PROCESSED:
        return O.g1_from_compressed(b)
    return O.g1_from_uncompressed(b, check_subgroup=fmt == mzk.SERDE_RAW_BYTES)


def sample_limbs(n, seed):
    """n affine points [s_i] G (Montgomery limbs) with identities at a few places."""
    pts = C.g1_mul_generator_many(C.fr_from_seed(seed, 3, n))
    for i in (0, n // 2, n - 1):
        pts[i] = 0
    return pts


@pytest.mark.parametrize("fmt", FORMATS)
def test_encode_matches_oracle(ctx, fmt):
    limbs = sample_limbs(300, 1)
    srs = ctx.srs_upload(limbs)
    got = srs.write_points(fmt)
    pb = mzk.point_bytes(fmt)
    assert len(got) == 300 * pb
    for i in range(300):
        P = O.g1_affine_from_mont_limbs(limbs[i])
        assert got[i * pb:(i + 1) * pb] == encode(P, fmt), i
    srs.close()


@pytest.mark.parametrize("fmt", FORMATS)
def test_decode_matches_oracle(ctx, fmt):
    limbs = sample_limbs(200, 2)
    pts = [O.g1_affine_from_mont_limbs(r) for r in limbs]
    pts += [O.g1_neg(P) for P in pts[:50]]
    data = b"".join(encode(P, fmt) for P in pts)
    assert [decode(data[i * mzk.point_bytes(fmt):(i + 1) * mzk.point_bytes(fmt)], fmt)
            for i in range(len(pts))] == pts
    srs = mzk.ResidentSrs.read_points(ctx, data, len(pts), fmt)
    want = np.array([O.g1_affine_mont_limbs(P) for P in pts], dtype=np.uint64)
    assert np.array_equal(srs.download(), want)
    srs.close()


@pytest.mark.parametrize("fmt", [mzk.SERDE_PROCESSED, mzk.SERDE_RAW_BYTES])
def test_round_trip_across_chunks(ctx, fmt):
    n = (1 << 18) + 5          # more than one staged chunk
    tau = 0xC0FFEE1234567 % O.R_MOD
    g = ctx.srs_generate(plonk.mont_limbs(tau), n)
    data = g.write_points(fmt)
    back = mzk.ResidentSrs.read_points(ctx, data, n, fmt)
    assert np.array_equal(back.download(), g.download())
    for i in (0, 1, (1 << 18) - 1, 1 << 18, n - 1):
        P = O.srs_point(tau, i)
        pb = mzk.point_bytes(fmt)
        assert data[i * pb:(i + 1) * pb] == encode(P, fmt)
    g.close()
    back.close()


def non_g1_point():
    x = 5
    while True:
        rhs = (x ** 3 + 4) % O.P_MOD
        y = pow(rhs, (O.P_MOD + 1) // 4, O.P_MOD)
        if y * y % O.P_MOD == rhs and not O.g1_in_subgroup((x, y)):
            return (x, y)
        x += 1


def bad_records(fmt):
    """(bytes, reason) for one bad point of each class the format checks."""
    P = O.g1_mul(O.G1_GEN, 777)
    good = bytearray(encode(P, fmt))
    out = []
    flip = bytearray(good)
    flip[0] ^= 0x80                               # wrong compression flag
    out.append((bytes(flip), "invalid encoding"))
    big = bytearray(O.P_MOD.to_bytes(48, "big")) + (good[48:] if fmt != mzk.SERDE_PROCESSED else b"")
    if fmt == mzk.SERDE_PROCESSED:
        big[0] |= 0x80
    out.append((bytes(big), "invalid encoding"))
    ident = bytearray(encode(None, fmt))
    ident[-1] = 1                                 # identity flag with a non-zero coordinate
    out.append((bytes(ident), "invalid encoding"))
    if fmt == mzk.SERDE_PROCESSED:
        x = 1
        while pow((x ** 3 + 4) % O.P_MOD, (O.P_MOD - 1) // 2, O.P_MOD) == 1:
            x += 1
        nb = bytearray(x.to_bytes(48, "big"))
        nb[0] |= 0x80
        out.append((bytes(nb), "not on the curve"))
    else:
        out.append((O.g1_to_uncompressed((P[0], (P[1] + 1) % O.P_MOD)), "not on the curve"))
        sgn = bytearray(good)
        sgn[0] |= 0x20                            # sign flag in an uncompressed point
        out.append((bytes(sgn), "invalid encoding"))
    if fmt != mzk.SERDE_RAW_BYTES_UNCHECKED:
        out.append((encode((0, 2), fmt), "not in the prime-order subgroup"))
        out.append((encode(non_g1_point(), fmt), "not in the prime-order subgroup"))
    return out


@pytest.mark.parametrize("fmt", FORMATS)
def test_rejections_name_lowest_index(ctx, fmt):
    limbs = sample_limbs(100, 3)
    pts = [O.g1_affine_from_mont_limbs(r) for r in limbs]
    recs = [encode(P, fmt) for P in pts]
    for bad, why in bad_records(fmt):
        for idx in (0, 37, 99):
            data = list(recs)
            data[idx] = bad
            if idx < 99:
                data[99] = bad                    # a second bad point: the lower index is named
            with pytest.raises(mzk.MzkError) as e:
                mzk.ResidentSrs.read_points(ctx, b"".join(data), 100, fmt)
            assert f"point {idx}: {why}" in str(e.value), (why, idx, str(e.value))


def test_unchecked_accepts_points_outside_g1(ctx):
    fmt = mzk.SERDE_RAW_BYTES_UNCHECKED
    Q = non_g1_point()
    pts = [O.g1_mul(O.G1_GEN, 3), (0, 2), Q, None]
    srs = mzk.ResidentSrs.read_points(ctx, b"".join(encode(P, fmt) for P in pts), 4, fmt)
    assert np.array_equal(srs.download(), np.array([O.g1_affine_mont_limbs(P) for P in pts], dtype=np.uint64))
    srs.close()


@pytest.mark.parametrize("fmt", FORMATS)
def test_params_file_round_trip(ctx, fmt):
    k = 10
    tau = 0xABCDEF987654321 % O.R_MOD
    g = ctx.srs_generate(plonk.mont_limbs(tau), 1 << k)
    gl = g.to_lagrange(k)
    g2len = mzk.point_bytes(fmt, g2=True)
    g2 = bytes(range(g2len % 256)) + bytes(g2len - g2len % 256)
    s_g2 = bytes(reversed(g2))
    blob = k.to_bytes(4, "little") + g.write_points(fmt) + gl.write_points(fmt) + g2 + s_g2
    p = mzk.KzgParams.read_custom(ctx, blob, fmt)
    assert p.k == k and p.g2 == g2 and p.s_g2 == s_g2
    assert np.array_equal(p.g.download(), g.download())
    assert np.array_equal(p.g_lagrange.download(), gl.download())
    d1 = gl.suffix_sum()
    assert np.array_equal(p.g_lagrange_delta.download(), d1.download())
    assert np.array_equal(p.g_lagrange_double_delta.download(), d1.suffix_sum().download())
    assert p.write_custom(fmt) == blob
    with pytest.raises(ValueError):
        mzk.KzgParams.read_custom(ctx, blob[:-1], fmt)
