"""The C ABI library: it loads, exports every entry point include/mzk_gpu.h
declares, and reports errors as status codes (CPU only: no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.helpers import ROOT, mzk

HEADER = os.path.join(ROOT, "include", "mzk_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|size_t|char)\s*\*?\s*(mzk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "mzk_msm_best" in names and "mzk_msm_specific" in names
    assert "mzk_best_fft" in names and "mzk_coeff_to_extended_dev" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(mzk.LIB_PATH), "libmzk_gpu.so is not built"
    out = subprocess.run(["nm", "-D", "--defined-only", mzk.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(mzk_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_python_mirror_binds_every_symbol():
    L = mzk.load_library()
    assert L.mzk_version() >= 100
    assert set(mzk.SIGNATURES) >= set(declared_functions())


def test_errors_are_status_codes_not_aborts():
    L = mzk.load_library()
    # null out-pointer: rejected before touching any device
    assert L.mzk_ctx_create(0, None) == 1
    assert b"null" in L.mzk_last_error()
    # null context
    assert L.mzk_ctx_synchronize(None) == 1
    out = (ctypes.c_uint64 * 18)()
    assert L.mzk_msm_best(None, None, None, 4, out) == 1
    assert L.mzk_domain_info(None, None, None) == 1


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(mzk.MzkError):
        mzk.Context(0)
