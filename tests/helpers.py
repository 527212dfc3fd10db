"""Shared helpers for the tests: the product package, device transfers, and
conversions between limb arrays and the oracle's integers."""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

pkg = importlib.import_module("midnight-zk_amd")
mzk = pkg.mzk


def to_dev(a):
    import torch
    arr = np.ascontiguousarray(a, dtype=np.uint64)
    return torch.from_numpy(arr.view(np.int64)).cuda()


def from_dev(t):
    return t.cpu().numpy().view(np.uint64)


def empty_dev(rows, cols):
    import torch
    return torch.zeros((rows, cols), dtype=torch.int64, device="cuda")


# Host builds of the native check programs: plain (-O2), and with AddressSanitizer and
# UndefinedBehaviorSanitizer (any report aborts the program, so the test fails).
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=all"]
BUILD_MODES = ["plain", "asan_ubsan"]


def build_native(compiler: str, src: str, out: str, mode: str, lang_flags=("-std=c++17", "-x", "c++"),
                 libs=()):
    import subprocess
    opt = SANITIZE_FLAGS if mode == "asan_ubsan" else ["-O2"]
    subprocess.run([compiler, *opt, *lang_flags, src, "-o", out, *libs], check=True)
    return out
