"""MI355X (gfx950) prover hot path for midnight-zk: BLS12-381 G1 MSM and Fr NTTs.

The directory name is not an identifier; import it with
``importlib.import_module("midnight-zk_amd")``.  ``mzk`` holds the host mirror of
the reference's MSM / NTT interface over the C ABI in libmzk_gpu.so.
"""
from . import mzk  # noqa: F401
from . import plonk  # noqa: F401
from . import serve  # noqa: F401
from .mzk import (Context, EvaluationDomain, MzkError, ResidentSrs,  # noqa: F401
                  load_library)
