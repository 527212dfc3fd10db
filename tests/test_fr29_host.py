"""The radix-2^29 scalar field of the NTT passes, compiled for the host with g++ and
checked against Python integers: conversions both ways, Montgomery products, the small
reduction, lazy sums and differences, and eleven chained butterfly stages from the
largest inputs the NTT feeds them (CPU only)."""
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from tests.helpers import BUILD_MODES, build_native

from oracle import bls12_381 as O

HERE = os.path.dirname(os.path.abspath(__file__))
R = O.R_MOD
RINV = pow(1 << 256, -1, R)


@pytest.fixture(scope="module", params=BUILD_MODES)
def exe(tmp_path_factory, request):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fr29_" + request.param) / "fr29_check")
    return build_native(gxx, os.path.join(HERE, "native", "fr29_check.cpp"), out, request.param)


def words(v):
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def run(exe, op, xs, ys, out_words=8):
    head = f"{op} {len(xs)}\n".encode()
    payload = np.array([w for v in xs for w in words(v)], dtype="<u4").tobytes()
    payload += np.array([w for v in ys for w in words(v)], dtype="<u4").tobytes()
    res = subprocess.run([exe], input=head + payload, capture_output=True, check=True).stdout
    arr = np.frombuffer(res, dtype="<u4").reshape(len(xs), out_words)
    return [[sum(int(arr[i, 8 * h + j]) << (32 * j) for j in range(8)) for h in range(out_words // 8)]
            for i in range(len(xs))]


def values(rng, n):
    edge = [0, 1, 2, R - 1, R - 2, (R - 1) // 2, 1 << 254, R - (1 << 200)]
    return edge + [rng.randrange(R) for _ in range(n - len(edge))]


def test_conversions_and_products(exe):
    rng = random.Random(29)
    xs, ys = values(rng, 400), values(rng, 400)[::-1]
    for (got,), x in zip(run(exe, 0, xs, ys), xs):
        assert got == x
    for (got,), x in zip(run(exe, 2, xs, ys), xs):
        assert got == x
    for (got,), x, y in zip(run(exe, 1, xs, ys), xs, ys):
        assert got == x * y * RINV % R          # the Montgomery product of two blst values
    for (got,), x, y in zip(run(exe, 3, xs, ys), xs, ys):
        assert got == (x + y * x * RINV) % R
    for (got,), x, y in zip(run(exe, 4, xs, ys), xs, ys):
        assert got == (x - y * x * RINV) % R


def test_chained_butterflies(exe):
    rng = random.Random(30)
    xs, ys = values(rng, 300), values(rng, 300)
    for (gx, gy), x, y in zip(run(exe, 5, xs, ys, 16), xs, ys):
        X, Y = x, y                               # std values x 2^256; w is y's value
        for _ in range(11):
            t = y * Y * RINV % R
            X, Y = (X + t) % R, (X - t) % R
        assert (gx, gy) == (X, Y)


def test_paired_stages_with_lazy_sums(exe):
    """The NTT's register pair of stages with the first stage's sums left unnormalised
    (limbs up to 2^31 into the second stage's products), from inputs raised by 22 r."""
    rng = random.Random(31)
    xs, ys = values(rng, 300), values(rng, 300)[::-1]
    xs += [R - 1] * 4 + [(1 << 254) + i for i in range(4)]
    ys += [R - 1, R - 2, 1, 0] + [R - 1 - i for i in range(4)]
    for outs, a, b in zip(run(exe, 6, xs, ys, 32), xs, ys):
        x0, x1, x2, x3 = a, b, b, a                 # std values x 2^256; twiddles b, a
        t1, t3 = x1 * b * RINV % R, x3 * b * RINV % R
        y0, y1, y2, y3 = x0 + t1, x0 - t1, x2 + t3, x2 - t3
        t2, t4 = y2 * a * RINV % R, y3 * a * RINV % R
        assert outs == [(y0 + t2) % R, (y0 - t2) % R, (y1 + t4) % R, (y1 - t4) % R]
