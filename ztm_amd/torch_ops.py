"""Differentiable gravity for torch: thin Python face of the libtorch
extension ztm_amd._ztm_torch (ztm_amd/csrc/gfm_torch.cpp).

    g    = gravity(rho, plan)          # G rho, backward G^T grad (adjoint kernel)
    Y    = gravity_batched(P, plan)    # G P for P [B, N], backward batched G^T
    s    = smoothness(rho, plan)       # 0.5 ||grad rho||^2, backward L^T L rho
    loss = misfit(rho, plan, d, lam)   # 0.5 ||G rho - d||^2 + lam * s

rho is a float32 (FP32 path) or float64 (FP64 path) CUDA tensor on the plan's
device; g is float64. The extension is required: there is no Python fallback.
"""
from __future__ import annotations

import importlib

from .gfm import Plan

_ext = None


def ext():
    global _ext
    if _ext is None:
        try:
            _ext = importlib.import_module("ztm_amd._ztm_torch")
        except ImportError as e:   # pragma: no cover - exercised only when unbuilt
            raise ImportError("ztm_amd._ztm_torch is not built; run `python -m ztm_amd.build`"
                              ) from e
    return _ext


def gravity(rho, plan: Plan):
    return ext().gravity(rho, plan.handle)


def gravity_batched(rho, plan: Plan):
    return ext().gravity_batched(rho, plan.handle)


def smoothness(rho, plan: Plan):
    return ext().smoothness(rho, plan.handle)


def adjoint(r, plan: Plan):
    return ext().adjoint(r, plan.handle)


def misfit(rho, plan: Plan, d, lam: float):
    """0.5 ||G rho - d||^2 + 0.5 lam ||grad rho||^2 (a 0-d float64 tensor)."""
    res = gravity(rho, plan) - d
    return 0.5 * (res * res).sum() + lam * smoothness(rho, plan)
