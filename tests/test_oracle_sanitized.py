"""The oracle's C restatement (oracle/cpu_ref.c) compiled into a driver that runs every
entry point on seeded inputs and cross-checks the results, built plain and with
AddressSanitizer + UndefinedBehaviorSanitizer (any report aborts, so the run fails)
(CPU only)."""
import os
import shutil
import subprocess

import pytest

from tests.helpers import BUILD_MODES, build_native

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("mode", BUILD_MODES)
def test_cpu_ref_entry_points_clean(tmp_path, mode):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = build_native(gcc, os.path.join(HERE, "native", "cpu_ref_check.c"), str(tmp_path / "cpu_ref_check"),
                       mode, lang_flags=("-std=gnu11", "-pthread"), libs=("-lm",))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-4000:]
