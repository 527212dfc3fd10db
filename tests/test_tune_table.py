"""The autotuner's plan table (host side, CPU only): export / import round trip,
comment lines, all-or-nothing rejection of malformed text, reset."""
import pytest

from tests.helpers import mzk


@pytest.fixture(autouse=True)
def clean_table():
    mzk.tune_reset()
    yield
    mzk.tune_reset()


def test_empty_table_exports_nothing():
    assert mzk.tune_export() == ""


def test_round_trip():
    text = "msm 12 0 4 256 0\nntt 14 6 4096\nmsm 20 1 0 0 1024\nntt 22 11 256\n"   # export order: by size
    mzk.tune_import("# tuned on one box\n" + text)
    assert mzk.tune_export() == text
    mzk.tune_import(mzk.tune_export())
    assert mzk.tune_export() == text


@pytest.mark.parametrize("bad", [
    "msm 12 0 5 256 0",        # window bits other than 4, 8, 16
    "msm 40 0 4 256 0",        # size out of range
    "msm 12 2 4 256 0",        # unknown path
    "ntt 14 12 256",           # sub-DFT above 2^11
    "ntt 14 6",                # missing field
    "fft 14 6 256",            # unknown entry
    "msm 12 0 4 256 0 9",      # trailing field
])
def test_rejects_malformed(bad):
    mzk.tune_import("ntt 16 8 1024\n")
    with pytest.raises(mzk.MzkError):
        mzk.tune_import("msm 10 0 8 512 0\n" + bad + "\n")
    assert mzk.tune_export() == "ntt 16 8 1024\n"      # nothing of the failed text applied


def test_reset():
    mzk.tune_import("msm 12 0 4 256 0\n")
    mzk.tune_reset()
    assert mzk.tune_export() == ""
