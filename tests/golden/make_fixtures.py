"""Writes tests/golden/fixtures.json: digests of the NTT suite and the SRS-commit MSMs on
seeded inputs, for the parity tests to check the device (and the C restatement) against
without recomputing them.

  * NTT: forward transforms at 2^16, 2^20, 2^22 of the seeded stream (seed 0x77, stream 2,
    the device's mzk_fr_from_seed_dev): SHA-256 of the output limbs (Montgomery, LE u64),
    and sampled outputs re-derived independently by Horner evaluation of the input
    polynomial at omega^j (pure Python integers).
  * MSM: the commitment of the benchmark configuration (tau = stream (0x5EED, 0) element 0,
    scalars = stream (0x5EED, 1), bases [tau^i] G) at n = 2^16 and 2^20, from the closed
    form [sum_i s_i tau^i] G, as compressed bytes.

Run from the repository root: python tests/golden/make_fixtures.py (about a minute).
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bls12_381 as O  # noqa: E402
from oracle import cpu_ref as C  # noqa: E402

NTT_SEED, NTT_STREAM = 0x77, 2
MSM_SEED = 0x5EED


def sample_indices(n):
    return sorted({0, 1, 2, n // 3, n // 2, n - 1, (0x9E3779B9 * 7) % n})


def main():
    C.set_threads(os.cpu_count() or 8)
    out = {"ntt": [], "msm": []}
    for log_n in (16, 20, 22):
        n = 1 << log_n
        a = C.fr_from_seed(NTT_SEED, NTT_STREAM, n)
        y = C.ntt(a, log_n)
        coeffs = C.fr_rows_to_ints(a)
        w = O.root_of_unity(log_n)
        idx = sample_indices(n)
        vals = C.fr_rows_to_ints(y[idx])
        for j, v in zip(idx, vals):
            assert O.poly_eval(coeffs, pow(w, j, O.R_MOD)) == v, (log_n, j)
        out["ntt"].append({"log_n": log_n, "seed": NTT_SEED, "stream": NTT_STREAM,
                           "sha256": hashlib.sha256(y.astype("<u8").tobytes()).hexdigest(),
                           "samples": {str(j): hex(v) for j, v in zip(idx, vals)}})
        print("ntt", log_n, "ok", flush=True)
    tau = C.fr_rows_to_ints(C.fr_from_seed(MSM_SEED, 0, 1))[0]
    for log_n in (16, 20):
        n = 1 << log_n
        s = C.fr_rows_to_ints(C.fr_from_seed(MSM_SEED, 1, n))
        P = O.msm_srs_closed_form(s, tau)
        out["msm"].append({"log_n": log_n, "seed": MSM_SEED, "tau_stream": 0, "scalar_stream": 1,
                           "compressed": O.g1_to_compressed(P).hex()})
        print("msm", log_n, "ok", flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
