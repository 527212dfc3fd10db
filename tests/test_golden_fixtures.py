"""The committed fixtures (tests/golden/fixtures.json, written by make_fixtures.py) against
the oracle: the C restatement's NTTs reproduce the output digests, sampled outputs equal
Horner evaluations of the input at omega^j, and the closed-form commitment reproduces the
MSM points (CPU only)."""
import hashlib
import json
import os

import pytest

from oracle import bls12_381 as O
from oracle import cpu_ref as C

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


@pytest.mark.parametrize("entry", FIX["ntt"], ids=lambda e: f"2^{e['log_n']}")
def test_ntt_digest(entry):
    C.set_threads(8)
    log_n = entry["log_n"]
    a = C.fr_from_seed(entry["seed"], entry["stream"], 1 << log_n)
    y = C.ntt(a, log_n)
    assert hashlib.sha256(y.astype("<u8").tobytes()).hexdigest() == entry["sha256"]
    idx = [int(j) for j in entry["samples"]]
    got = C.fr_rows_to_ints(y[idx])
    assert got == [int(entry["samples"][str(j)], 16) for j in idx]
    if log_n <= 16:   # an independent closed form for the samples (pure Python Horner)
        coeffs = C.fr_rows_to_ints(a)
        w = O.root_of_unity(log_n)
        for j, v in zip(idx, got):
            assert O.poly_eval(coeffs, pow(w, j, O.R_MOD)) == v


def test_msm_closed_form_2p16():
    entry = next(e for e in FIX["msm"] if e["log_n"] == 16)
    tau = C.fr_rows_to_ints(C.fr_from_seed(entry["seed"], entry["tau_stream"], 1))[0]
    s = C.fr_rows_to_ints(C.fr_from_seed(entry["seed"], entry["scalar_stream"], 1 << 16))
    assert O.g1_to_compressed(O.msm_srs_closed_form(s, tau)).hex() == entry["compressed"]
