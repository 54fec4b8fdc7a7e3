"""ztm_amd -- MI355X (gfx950) gravity forward/adjoint engine for abatchev/ZTM.

The compute lives in libztm_gfm.so (hand-written HIP kernels behind the C ABI
in include/ztm_gfm.h); this package holds its Python host mirror (gfm), the
seeded synthetic workloads (workloads) and the in-tree build (build).
"""
from . import gfm, workloads  # noqa: F401
from .gfm import GF, FP32, FP64, Plan, ZtmError, gfm_host  # noqa: F401

__all__ = ["gfm", "workloads", "GF", "gfm_host", "Plan", "ZtmError", "FP32", "FP64"]
