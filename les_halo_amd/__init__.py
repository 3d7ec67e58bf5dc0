"""LES_Halo time step on AMD MI355X (gfx950): HIP kernels behind a C ABI (libles_hip.so),
with a Python mirror of the reference's GPU host API."""
from . import fields  # noqa: F401
from .les import (LES, Ensemble, LESError, LES_PREC_FP32, LES_PREC_MIXED,  # noqa: F401
                  LES_SUM_EXACT, LES_SUM_REFERENCE, TiledLES)
