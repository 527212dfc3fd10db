"""The prover's host BLAKE2b (keyed transcript hash and personalised verifying-key
hash), compiled with g++ and compared with hashlib.blake2b on messages of every length
around the 128-byte block boundary, with and without key and personalisation (CPU only).
hashlib implements RFC 7693, whose test vector is checked first."""
import hashlib
import os
import random
import shutil
import subprocess

import pytest

from tests.helpers import BUILD_MODES, build_native

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", params=BUILD_MODES)
def exe(tmp_path_factory, request):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("b2b_" + request.param) / "blake2b_check")
    return build_native(gxx, os.path.join(HERE, "native", "blake2b_check.cpp"), out, request.param)


def digests(exe, cases):
    payload = b"".join(f"{len(k)} {len(p)} {len(m)}\n".encode() + k + p + m for k, p, m in cases)
    r = subprocess.run([exe], input=payload, capture_output=True, check=True)
    out = r.stdout
    assert len(out) == 64 * len(cases)
    return [out[64 * i:64 * i + 64] for i in range(len(cases))]


def test_rfc7693_vector(exe):
    # RFC 7693 appendix A: BLAKE2b-512("abc")
    want = bytes.fromhex(
        "ba80a53f981c4d0d6a2797b69f12f6e94c212f14685ac4b74b12bb6fdbffa2d1"
        "7d87c5392aab792dc252d5de4533cc9518d38aa8dbf1925ab92386edd4009923")
    assert digests(exe, [(b"", b"", b"abc")])[0] == want


def test_against_hashlib(exe):
    rng = random.Random(11)
    cases = []
    for n in list(range(0, 300)) + [1000, 4096, 10000]:
        msg = bytes(rng.getrandbits(8) for _ in range(n))
        key = b"Domain separator for transcript" if n % 3 == 0 else bytes(rng.getrandbits(8) for _ in range(n % 65))
        pers = b"Halo2-Verify-Key" if n % 2 else b""
        cases.append((key, pers, msg))
    got = digests(exe, cases)
    for (key, pers, msg), d in zip(cases, got):
        assert d == hashlib.blake2b(msg, key=key, person=pers).digest(), (len(key), len(pers), len(msg))
