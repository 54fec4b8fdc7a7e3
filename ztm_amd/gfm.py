"""Python host mirror of libztm_gfm's C ABI (include/ztm_gfm.h).

The reference's gravity path is called three ways, and each has a mirror here:

  GF(bfr, di, dj, dk, fh, fv, g, rho, s)
      the CUDA kernel's call (GFM_GPU.cu:16, launch :31) on device tensors;
      g is accumulated into at interior stations.
  gfm_host(bfr, di, dj, dk, fh, fv, rho, s) -> g
      the GFM program body (GFM_GPU.cu:24-34) on host float32 arrays.
  Plan(grid, stations, ...).forward(rho) / .adjoint(r)
      the CPU gravity lambda (TFM.cpp:95-99) generalised to explicit geometry,
      FP32/FP64 and its transpose.

Everything runs through libztm_gfm.so; if it is missing or fails to load this
module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libztm_gfm.so")

FP32, FP64, FP64_ZONED = 0, 1, 2
POINT_MASS, PRISM = 0, 1
KAPPA_SI = 6.674e-6
KAPPA_REF = 66.7
_ERRNAMES = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ENODEV", 5: "EIO"}


class ZtmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ztm {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class CGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class CProblem(ctypes.Structure):
    _fields_ = [("grid", CGrid), ("formula", ctypes.c_int32), ("kappa", ctypes.c_double),
                ("threshold", ctypes.c_double)]


# name -> (restype, argtypes); every symbol declared in include/ztm_gfm.h.
_P, _I, _I64, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "ztm_version": (_I, []),
    "ztm_strerror": (ctypes.c_char_p, [_I]),
    "ztm_last_error": (ctypes.c_char_p, []),
    "ztm_init": (_I, []),
    "ztm_gf": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P]),
    "ztm_gf_async": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "ztm_gf_noalloc": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "ztm_gfm_host": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _I64, _P]),
    "ztm_plan_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(CProblem), _I64, _P]),
    "ztm_plan_retain": (_I, [_P]),
    "ztm_plan_destroy": (_I, [_P]),
    "ztm_plan_num_cells": (_I64, [_P]),
    "ztm_plan_num_stations": (_I64, [_P]),
    "ztm_plan_get_problem": (_I, [_P, ctypes.POINTER(CProblem)]),
    "ztm_plan_forward": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_adjoint": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_smoothness": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_forward_batched": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_adjoint_batched": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_forward_batched64": (_I, [_P, _I, _P, _P, _P]),
    "ztm_calib_mfma": (_I, [_I, _I, _I, _DP, _P]),
    "ztm_gn_workspace_bytes": (_I64, [_I64, _I, _I]),
    "ztm_gn_normal": (_I, [_I64, _I, _P, _P, _P, _P, _P, _P]),
    "ztm_gn_damped_steps": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ztm_gn_inverse_diag": (_I, [_I, _P, _D, _P, _P, _P]),
    "ztm_plan_adjoint_batched64": (_I, [_P, _I, _P, _P, _P]),
    "ztm_plan_set_batched_tuning": (_I, [_P, _I, _I]),
    "ztm_plan_get_batched_tuning": (_I, [_P, _IP, _IP, _IP, _IP]),
    "ztm_plan_set_batched_variant": (_I, [_P, _I]),
    "ztm_plan_set_batched_kernel": (_I, [_P, _I, _I, _I, _I, _I]),
    "ztm_plan_get_batched_kernel": (_I, [_P, _I, _I, _IP, _IP, _IP]),
    "ztm_plan_autotune_batched": (_I, [_P, _I, _I, _P]),
    "ztm_plan_store_bytes": (_I64, [_P]),
    "ztm_plan_store": (_I, [_P, _P]),
    "ztm_plan_store_release": (_I, [_P]),
    "ztm_plan_forward_stored": (_I, [_P, _P, _P, _P]),
    "ztm_plan_adjoint_stored": (_I, [_P, _P, _P, _P]),
    "ztm_plan_set_tuning": (_I, [_P, _I, _I, _I, _I]),
    "ztm_plan_get_tuning": (_I, [_P, _IP, _IP, _IP, _IP]),
    "ztm_tfm_drho": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P, _P]),
    "ztm_tfm_gravity_host": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _D, _D, _D, _P,
                                  _P]),
    "ztm_tfm_gravity_batched": (_I, [_I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _I64, _P, _I64,
                                     _P, _P, _P, _P, _P]),
    "ztm_tfm_drho_jvp": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P, _P]),
    "ztm_tfm_gravity_jvp": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P, _P,
                                 _P, _P]),
    "ztm_tfm_gravity_vjp": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P, _P,
                                 _P, _P, _P]),
    "ztm_tfm_gravity_jvp_grids": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P,
                                       _I, _P, _P, _P, _P, _P]),
    "ztm_tfm_gravity_vjp_grids": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _P,
                                       _P, _P, _P, _P]),
    "ztm_batch_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(CProblem),
                              ctypes.POINTER(_I64), ctypes.POINTER(_P)]),
    "ztm_batch_destroy": (_I, [_P]),
    "ztm_batch_num_problems": (_I, [_P]),
    "ztm_batch_layout": (_I, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "ztm_batch_forward": (_I, [_P, _P, _P, _P]),
    "ztm_batch_adjoint": (_I, [_P, _P, _P, _P]),
    "ztm_batch_solve": (_I, [_P, _P, ctypes.POINTER(_D), _I, _P, _P, _I, _P]),
    "ztm_batch_set_tuning": (_I, [_P, _I, _I, _I, _I, _I]),
    "ztm_batch_get_tuning": (_I, [_P, _I, _IP, _IP, _IP, _IP, _IP]),
    "ztm_batch_autotune": (_I, [_P, _I, _P]),
    "ztm_profile_enable": (_I, [_I]),
    "ztm_profile_read": (_I, [ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "ztm_profile_reset": (_I, []),
}

_lib = None


def lib():
    """Load libztm_gfm.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m ztm_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise ZtmError(rc, lib().ztm_last_error().decode())


def _stream(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _ptr(t) -> int:
    return t.data_ptr()


# ------------------------------------------------------- reference mirrors
def GF(bfr, di, dj, dk, fh, fv, g, rho, s, stream=None) -> None:
    """GFM_GPU.cu:16 on device float32 tensors; g += interior-station gravity."""
    import torch
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    for name, t, size in (("g", g, di * dj), ("rho", rho, n), ("s", s, di * dj)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32
                and t.is_contiguous() and t.numel() == size):
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of {size} values")
    check(lib().ztm_gf_async(bfr, di, dj, dk, fh, fv, _ptr(g), _ptr(rho), _ptr(s),
                             _stream(stream)))


def gfm_host(bfr, di, dj, dk, fh, fv, rho, s) -> np.ndarray:
    """GFM_GPU.cu:24-34 on host float32 arrays; returns g[di*dj] (float32)."""
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    s = np.ascontiguousarray(s, dtype=np.float32)
    if s.size != di * dj:
        raise ValueError(f"s needs di*dj = {di * dj} values")
    g = np.zeros(di * dj, dtype=np.float32)
    check(lib().ztm_gfm_host(bfr, di, dj, dk, fh, fv, g.ctypes.data, rho.ctypes.data, rho.size,
                             s.ctypes.data))
    return g


# ------------------------------------------ gravity producer (TFM.cpp:83-94)
def _host_f64(a, n: int, name: str) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"{name} needs {n} values, got {a.size}")
    return a


def tfm_drho(di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, out=None, stream=None):
    """TFM.cpp:83-89 on the device: sea/moho float64 CUDA tensors [di*dj], vel
    [di*dj*dk]; returns drho float32 [ii*jj*kk] (the reference's value, narrowed
    as its GPU path does)."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    for name, t, size in (("sea", sea, di * dj), ("moho", moho, di * dj),
                          ("vel", vel, di * dj * dk)):
        _check_t(t, torch.float64, size, name)
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=sea.device)
    _check_t(out, torch.float32, n, "drho")
    check(lib().ztm_tfm_drho(di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea), _ptr(moho),
                             _ptr(vel), vm, ro, dr, _ptr(out), _stream(stream)))
    return out


def tfm_gravity_host(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, surf, vm, ro, dr,
                     want_drho=False):
    """One gravity forward model of TFM.cpp:83-94 in process on host arrays:
    returns G[di*dj] (float64) and, if asked, drho (float32)."""
    rc = np.ascontiguousarray(rc, dtype=np.float64).ravel()
    if rc.size < 7:
        raise ValueError("rc needs RC[0..6]")
    sea = _host_f64(sea, di * dj, "sea")
    moho = _host_f64(moho, di * dj, "moho")
    vel = _host_f64(vel, di * dj * dk, "vel")
    surf = _host_f64(surf, di * dj, "surf")
    G = np.empty(di * dj)
    drho = None
    if want_drho:
        drho = np.empty(((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1),
                        dtype=np.float32)
    check(lib().ztm_tfm_gravity_host(bfr, di, dj, dk, fh, fv, rc.ctypes.data, sea.ctypes.data,
                                     moho.ctypes.data, vel.ctypes.data, surf.ctypes.data, vm, ro,
                                     dr, G.ctypes.data,
                                     drho.ctypes.data if drho is not None else None))
    return (G, drho) if want_drho else G


def tfm_gravity_batched(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, params, s, drho=None,
                        g=None, stream=None):
    """B FM variants in one call (FD Jacobian columns): sea float64 [di*dj];
    moho float64 [di*dj] (shared) or [B, di*dj]; vel [di*dj*dk] or
    [B, di*dj*dk]; params float64 [B, 3] = (vm, ro, dr) on the device; s
    float32 [di*dj]. Returns (g float64 [B, di*dj], drho float32 [B, N])."""
    import torch
    rc = np.ascontiguousarray(rc, dtype=np.float64).ravel()
    B = int(params.shape[0])
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    M = di * dj
    _check_t(params, torch.float64, 3 * B, "params")
    _check_t(sea, torch.float64, M, "sea")
    _check_t(s, torch.float32, M, "s")
    ms = M if moho.numel() == B * M and B > 1 else 0
    vs = M * dk if vel.numel() == B * M * dk and B > 1 else 0
    _check_t(moho, torch.float64, B * M if ms else M, "moho")
    _check_t(vel, torch.float64, B * M * dk if vs else M * dk, "vel")
    if drho is None:
        drho = torch.empty(B, n, dtype=torch.float32, device=sea.device)
    if g is None:
        g = torch.empty(B, M, dtype=torch.float64, device=sea.device)
    _check_t(drho, torch.float32, B * n, "drho")
    _check_t(g, torch.float64, B * M, "g")
    check(lib().ztm_tfm_gravity_batched(bfr, di, dj, dk, fh, fv, B, rc.ctypes.data, _ptr(sea),
                                        _ptr(moho), ms, _ptr(vel), vs, _ptr(params), _ptr(s),
                                        _ptr(drho), _ptr(g), _stream(stream)))
    return g, drho


def tfm_drho_jvp(di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, out=None, stream=None):
    """d drho / d(vm, ro, dr) at one parameter point on the device: returns
    float32 [3, N] (0 where the producer or GF drops the cell)."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    for name, t, size in (("sea", sea, di * dj), ("moho", moho, di * dj),
                          ("vel", vel, di * dj * dk)):
        _check_t(t, torch.float64, size, name)
    if out is None:
        out = torch.empty(3, n, dtype=torch.float32, device=sea.device)
    _check_t(out, torch.float32, 3 * n, "ddrho")
    check(lib().ztm_tfm_drho_jvp(di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea), _ptr(moho),
                                 _ptr(vel), vm, ro, dr, _ptr(out), _stream(stream)))
    return out


def tfm_gravity_jvp(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, s, ddrho=None,
                    dg=None, stream=None):
    """dG / d(vm, ro, dr) at the interior stations: sea/moho float64 [di*dj],
    vel [di*dj*dk], s float32 [di*dj] on the device. Returns (dg float64
    [3, di*dj], ddrho float32 [3, N])."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    M = di * dj
    for name, t, size in (("sea", sea, M), ("moho", moho, M), ("vel", vel, M * dk)):
        _check_t(t, torch.float64, size, name)
    _check_t(s, torch.float32, M, "s")
    if ddrho is None:
        ddrho = torch.empty(3, n, dtype=torch.float32, device=sea.device)
    if dg is None:
        dg = torch.empty(3, M, dtype=torch.float64, device=sea.device)
    _check_t(ddrho, torch.float32, 3 * n, "ddrho")
    _check_t(dg, torch.float64, 3 * M, "dg")
    check(lib().ztm_tfm_gravity_jvp(bfr, di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea),
                                    _ptr(moho), _ptr(vel), vm, ro, dr, _ptr(s), _ptr(ddrho),
                                    _ptr(dg), _stream(stream)))
    return dg, ddrho


def tfm_gravity_vjp(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, s, r, ddrho=None,
                    out=None, stream=None):
    """Gradient of r . G with respect to (vm, ro, dr) for a station residual r
    (float64 [di*dj] on the device; exterior entries ignored). Returns (out
    float64 [3], ddrho float32 [3, N])."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    M = di * dj
    for name, t, size in (("sea", sea, M), ("moho", moho, M), ("vel", vel, M * dk),
                          ("r", r, M)):
        _check_t(t, torch.float64, size, name)
    _check_t(s, torch.float32, M, "s")
    if ddrho is None:
        ddrho = torch.empty(3, n, dtype=torch.float32, device=sea.device)
    if out is None:
        out = torch.empty(3, dtype=torch.float64, device=sea.device)
    _check_t(ddrho, torch.float32, 3 * n, "ddrho")
    _check_t(out, torch.float64, 3, "out")
    check(lib().ztm_tfm_gravity_vjp(bfr, di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea),
                                    _ptr(moho), _ptr(vel), vm, ro, dr, _ptr(s), _ptr(r),
                                    _ptr(ddrho), _ptr(out), _stream(stream)))
    return out, ddrho


def tfm_gravity_jvp_grids(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, s,
                          dmoho=None, dvel=None, ddrho=None, dg=None, stream=None):
    """dG along n_dirs tangent directions of the Moho grid (dmoho float64
    [n_dirs, di*dj]) and / or the velocity grid (dvel [n_dirs, di*dj*dk]) at
    one point; sea/moho [di*dj], vel [di*dj*dk] float64, s float32 [di*dj],
    all on the device. Returns (dg float64 [n_dirs, di*dj], ddrho float32
    [n_dirs, N])."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    M = di * dj
    for name, t, size in (("sea", sea, M), ("moho", moho, M), ("vel", vel, M * dk)):
        _check_t(t, torch.float64, size, name)
    _check_t(s, torch.float32, M, "s")
    if dmoho is None and dvel is None:
        raise ValueError("give dmoho and / or dvel")
    nd = int((dmoho if dmoho is not None else dvel).numel() //
             (M if dmoho is not None else M * dk))
    if dmoho is not None:
        _check_t(dmoho, torch.float64, nd * M, "dmoho")
    if dvel is not None:
        _check_t(dvel, torch.float64, nd * M * dk, "dvel")
    dev = sea.device
    if ddrho is None:
        ddrho = torch.empty(nd, n, dtype=torch.float32, device=dev)
    if dg is None:
        dg = torch.empty(nd, M, dtype=torch.float64, device=dev)
    _check_t(ddrho, torch.float32, nd * n, "ddrho")
    _check_t(dg, torch.float64, nd * M, "dg")
    check(lib().ztm_tfm_gravity_jvp_grids(
        bfr, di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea), _ptr(moho), _ptr(vel), vm, ro, dr,
        _ptr(s), nd, _ptr(dmoho) if dmoho is not None else None,
        _ptr(dvel) if dvel is not None else None, _ptr(ddrho), _ptr(dg), _stream(stream)))
    return dg, ddrho


def tfm_gravity_vjp_grids(bfr, di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr, s, r,
                          moho_grad=True, vel_grad=True, stream=None):
    """Gradients of r . G with respect to the Moho grid (float64 [di*dj]) and
    the velocity grid (float64 [di*dj*dk]) for a station residual r (float64
    [di*dj] on the device; exterior entries ignored). Returns (gmoho, gvel),
    None where not asked for."""
    import torch
    rc = _host_f64(rc, 7, "rc") if np.size(rc) == 7 else _host_f64(rc, 10, "rc")
    M = di * dj
    for name, t, size in (("sea", sea, M), ("moho", moho, M), ("vel", vel, M * dk),
                          ("r", r, M)):
        _check_t(t, torch.float64, size, name)
    _check_t(s, torch.float32, M, "s")
    gm = torch.empty(M, dtype=torch.float64, device=sea.device) if moho_grad else None
    gv = torch.empty(M * dk, dtype=torch.float64, device=sea.device) if vel_grad else None
    check(lib().ztm_tfm_gravity_vjp_grids(
        bfr, di, dj, dk, fh, fv, rc.ctypes.data, _ptr(sea), _ptr(moho), _ptr(vel), vm, ro, dr,
        _ptr(s), _ptr(r), _ptr(gm) if gm is not None else None,
        _ptr(gv) if gv is not None else None, _stream(stream)))
    return gm, gv


# ------------------------------------------------------------------ plans
@dataclass
class Tuning:
    st_per_lane: int = 0
    rows_per_slice: int = 0
    cells_per_lane: int = 0
    adj_splits: int = 0


class Plan:
    """One grid + station set on the current device (ztm_plan_*)."""

    def __init__(self, grid, stations, kappa=KAPPA_SI, threshold=0.0, formula=POINT_MASS):
        st = np.ascontiguousarray(stations, dtype=np.float64).reshape(-1, 3)
        prob = CProblem(CGrid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0,
                              grid.dx, grid.dy, grid.dz), formula, kappa, threshold)
        h = ctypes.c_void_p()
        check(lib().ztm_plan_create(ctypes.byref(h), ctypes.byref(prob), st.shape[0],
                                    st.ctypes.data if st.size else None))
        self._h = h
        self.grid = grid
        self.n_stations = st.shape[0]
        self.n_cells = grid.nx * grid.ny * grid.nz

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ztm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    def set_tuning(self, st_per_lane=0, rows_per_slice=0, cells_per_lane=0, adj_splits=0):
        check(lib().ztm_plan_set_tuning(self._h, st_per_lane, rows_per_slice, cells_per_lane,
                                        adj_splits))

    def tuning(self) -> Tuning:
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().ztm_plan_get_tuning(self._h, *[ctypes.byref(x) for x in v]))
        return Tuning(*[x.value for x in v])

    # raw device-pointer calls (stream-ordered)
    def forward_ptr(self, precision: int, rho_ptr: int, g_ptr: int, stream: int | None) -> None:
        check(lib().ztm_plan_forward(self._h, precision, rho_ptr, g_ptr, stream))

    def adjoint_ptr(self, precision: int, r_ptr: int, out_ptr: int, stream: int | None) -> None:
        check(lib().ztm_plan_adjoint(self._h, precision, r_ptr, out_ptr, stream))

    # torch-tensor conveniences
    def forward(self, rho, out=None, stream=None, zoned=False):
        """g[M] (float64) = kappa * sum_n rho[n] K(m,n); rho float32 -> FP32 path,
        float64 -> FP64 path (zoned=True: FP64 with the FP32 path's prism zones)."""
        import torch
        prec = _prec_of(rho, self.n_cells, "rho", zoned)
        if out is None:
            out = torch.empty(self.n_stations, dtype=torch.float64, device=rho.device)
        _check_t(out, torch.float64, self.n_stations, "g")
        self.forward_ptr(prec, _ptr(rho), _ptr(out), _stream(stream))
        return out

    # batched (ensemble) forward
    def forward_batched(self, rho, out=None, stream=None):
        """rho float32 [B, N] (B <= 256) -> g float64 [B, M] on split-BF16 MFMA."""
        import torch
        if not (isinstance(rho, torch.Tensor) and rho.is_cuda and rho.dtype == torch.float32
                and rho.is_contiguous() and rho.dim() == 2 and rho.shape[1] == self.n_cells):
            raise ValueError(f"rho must be a contiguous float32 CUDA tensor [B, {self.n_cells}]")
        B = rho.shape[0]
        if out is None:
            out = torch.empty(B, self.n_stations, dtype=torch.float64, device=rho.device)
        _check_t(out, torch.float64, B * self.n_stations, "g")
        check(lib().ztm_plan_forward_batched(self._h, B, _ptr(rho), _ptr(out), _stream(stream)))
        return out

    def adjoint_batched(self, r, out=None, stream=None):
        """r float32 [B, M] (B <= 256) -> G^T r float32 [B, N] on split-BF16 MFMA."""
        import torch
        if not (isinstance(r, torch.Tensor) and r.is_cuda and r.dtype == torch.float32
                and r.is_contiguous() and r.dim() == 2 and r.shape[1] == self.n_stations):
            raise ValueError(f"r must be a contiguous float32 CUDA tensor [B, {self.n_stations}]")
        B = r.shape[0]
        if out is None:
            out = torch.empty(B, self.n_cells, dtype=torch.float32, device=r.device)
        _check_t(out, torch.float32, B * self.n_cells, "out")
        check(lib().ztm_plan_adjoint_batched(self._h, B, _ptr(r), _ptr(out), _stream(stream)))
        return out

    def forward_batched64(self, rho, out=None, stream=None):
        """rho float64 [B, N] (B <= 256) -> g float64 [B, M]: the ensemble in FP64
        (G generated once per tile, reused across the batch on FP64 MFMA)."""
        import torch
        if not (isinstance(rho, torch.Tensor) and rho.is_cuda and rho.dtype == torch.float64
                and rho.is_contiguous() and rho.dim() == 2 and rho.shape[1] == self.n_cells):
            raise ValueError(f"rho must be a contiguous float64 CUDA tensor [B, {self.n_cells}]")
        B = rho.shape[0]
        if out is None:
            out = torch.empty(B, self.n_stations, dtype=torch.float64, device=rho.device)
        _check_t(out, torch.float64, B * self.n_stations, "g")
        check(lib().ztm_plan_forward_batched64(self._h, B, _ptr(rho), _ptr(out),
                                               _stream(stream)))
        return out

    def adjoint_batched64(self, r, out=None, stream=None):
        """r float64 [B, M] (B <= 256) -> G^T r float64 [B, N] on FP64 MFMA."""
        import torch
        if not (isinstance(r, torch.Tensor) and r.is_cuda and r.dtype == torch.float64
                and r.is_contiguous() and r.dim() == 2 and r.shape[1] == self.n_stations):
            raise ValueError(f"r must be a contiguous float64 CUDA tensor [B, {self.n_stations}]")
        B = r.shape[0]
        if out is None:
            out = torch.empty(B, self.n_cells, dtype=torch.float64, device=r.device)
        _check_t(out, torch.float64, B * self.n_cells, "out")
        check(lib().ztm_plan_adjoint_batched64(self._h, B, _ptr(r), _ptr(out), _stream(stream)))
        return out

    def batched_tuning(self):
        """(slices, forward split, adjoint split) in force for the batched calls."""
        return self.batched_settings()[:3]

    def batched_settings(self):
        """(slices, forward split, adjoint split, forward kernel variant)."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().ztm_plan_get_batched_tuning(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_batched_kernel(self, direction: int, n_models: int, variant: int, waves: int = 4,
                           slices: int = 0) -> None:
        """Kernel setting of the batched forward (direction 0) or adjoint (1)
        for the model-count class of n_models."""
        check(lib().ztm_plan_set_batched_kernel(self._h, direction, n_models, variant, waves,
                                                slices))

    def batched_kernel(self, direction: int, n_models: int):
        """(variant, waves per workgroup, K-slices) a batched call of n_models runs."""
        v = [ctypes.c_int() for _ in range(3)]
        check(lib().ztm_plan_get_batched_kernel(self._h, direction, n_models,
                                                *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune_batched(self, n_models: int, reps: int = 3, stream=None) -> None:
        """Time the batched kernel settings for n_models models and keep the
        fastest per direction (ztm_plan_autotune_batched)."""
        check(lib().ztm_plan_autotune_batched(self._h, n_models, reps, _stream(stream)))

    def set_batched_variant(self, variant: int) -> None:
        """Kernel variant (issue order) of the batched forward: 0 or 1."""
        check(lib().ztm_plan_set_batched_variant(self._h, variant))

    def set_batched_tuning(self, slices: int = 0, split: int = 0) -> None:
        """Forward K-slices (0 = choose) and split (2 = BF16 hi/lo, 3 = BF16
        three-way, 4 = FP16 hi/lo under power-of-two scales; 0 = keep); the
        split also applies to adjoint_batched."""
        check(lib().ztm_plan_set_batched_tuning(self._h, slices, split))

    # stored-matrix variant
    def store_bytes(self) -> int:
        return int(lib().ztm_plan_store_bytes(self._h))

    def store(self, stream=None) -> None:
        """Materialise G (FP32, entries evaluated in FP64) in device memory."""
        check(lib().ztm_plan_store(self._h, _stream(stream)))

    def store_release(self) -> None:
        check(lib().ztm_plan_store_release(self._h))

    def forward_stored(self, rho, out=None, stream=None):
        import torch
        _check_t(rho, torch.float32, self.n_cells, "rho")
        if out is None:
            out = torch.empty(self.n_stations, dtype=torch.float64, device=rho.device)
        _check_t(out, torch.float64, self.n_stations, "g")
        check(lib().ztm_plan_forward_stored(self._h, _ptr(rho), _ptr(out), _stream(stream)))
        return out

    def adjoint_stored(self, r, out=None, stream=None):
        import torch
        _check_t(r, torch.float32, self.n_stations, "r")
        if out is None:
            out = torch.empty(self.n_cells, dtype=torch.float32, device=r.device)
        _check_t(out, torch.float32, self.n_cells, "out")
        check(lib().ztm_plan_adjoint_stored(self._h, _ptr(r), _ptr(out), _stream(stream)))
        return out

    def smoothness(self, rho, out=None, stream=None):
        """out[N] = L^T L rho (gradient of 0.5 ||grad rho||^2), in rho's precision."""
        import torch
        prec = _prec_of(rho, self.n_cells, "rho")
        if out is None:
            out = torch.empty_like(rho)
        _check_t(out, rho.dtype, self.n_cells, "out")
        check(lib().ztm_plan_smoothness(self._h, prec, _ptr(rho), _ptr(out), _stream(stream)))
        return out

    def adjoint(self, r, out=None, stream=None, zoned=False):
        """out[N] = kappa * sum_m K(m,n) r[m], in r's precision (float32/float64;
        zoned=True as for forward)."""
        import torch
        prec = _prec_of(r, self.n_stations, "r", zoned)
        if out is None:
            out = torch.empty(self.n_cells, dtype=r.dtype, device=r.device)
        _check_t(out, r.dtype, self.n_cells, "out")
        self.adjoint_ptr(prec, _ptr(r), _ptr(out), _stream(stream))
        return out


# ---------------------------------------------------------- uneven batches
@dataclass
class BatchTuning:
    st_per_lane: int
    rows_per_slice: int
    cells_per_lane: int
    adj_splits: int
    size_class: int


class Batch:
    """Independent point-mass problems run together (ztm_batch_*): one
    persistent launch per forward/adjoint pass over all of them, and a batched
    CG solve of the regularised normal equations, optionally as a HIP graph.

    problems: sequence of (grid, stations[M,3], kappa) or (grid, stations,
    kappa, threshold). Device vectors are concatenated; cell_offsets /
    station_offsets give each problem's start (pack/unpack help)."""

    def __init__(self, problems):
        probs = []
        sts = []
        for pr in problems:
            grid, stations, kappa = pr[0], pr[1], pr[2]
            thr = pr[3] if len(pr) > 3 else 0.0
            st = np.ascontiguousarray(stations, dtype=np.float64).reshape(-1, 3)
            probs.append(CProblem(CGrid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0,
                                        grid.dx, grid.dy, grid.dz), POINT_MASS, kappa, thr))
            sts.append(st)
        n = len(probs)
        parr = (CProblem * n)(*probs)
        narr = (ctypes.c_int64 * n)(*[st.shape[0] for st in sts])
        sarr = (ctypes.c_void_p * n)(*[st.ctypes.data for st in sts])
        h = ctypes.c_void_p()
        check(lib().ztm_batch_create(ctypes.byref(h), n, parr, narr, sarr))
        self._h = h
        self.n_problems = n
        self.n_cells = [int(pr[0].nx * pr[0].ny * pr[0].nz) for pr in problems]
        self.n_stations = [st.shape[0] for st in sts]
        co = (ctypes.c_int64 * (n + 1))()
        so = (ctypes.c_int64 * (n + 1))()
        check(lib().ztm_batch_layout(h, co, so))
        self.cell_offsets = np.array(co[:], dtype=np.int64)
        self.station_offsets = np.array(so[:], dtype=np.int64)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ztm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def total_cells(self) -> int:
        return int(self.cell_offsets[-1])

    @property
    def total_stations(self) -> int:
        return int(self.station_offsets[-1])

    def pack_cells(self, arrays, dtype=None, device=None):
        """Concatenate per-problem cell vectors (padding = 0) into one tensor."""
        return self._pack(arrays, self.cell_offsets, self.n_cells, dtype, device)

    def pack_stations(self, arrays, dtype=None, device=None):
        return self._pack(arrays, self.station_offsets, self.n_stations, dtype, device)

    def unpack_cells(self, t):
        return [t[int(o):int(o) + n] for o, n in zip(self.cell_offsets, self.n_cells)]

    def unpack_stations(self, t):
        return [t[int(o):int(o) + n] for o, n in zip(self.station_offsets, self.n_stations)]

    @staticmethod
    def _pack(arrays, offs, sizes, dtype, device):
        import torch
        dtype = dtype or torch.float32
        out = torch.zeros(int(offs[-1]), dtype=dtype, device=device or "cuda")
        for a, o, n in zip(arrays, offs, sizes):
            a = torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a)
            if a.numel() != n:
                raise ValueError(f"expected {n} values, got {a.numel()}")
            out[int(o):int(o) + n] = a.reshape(-1).to(device=out.device, dtype=dtype)
        return out

    def forward(self, rho, out=None, stream=None):
        """rho float32 [total_cells] -> g float64 [total_stations]."""
        import torch
        _check_t(rho, torch.float32, self.total_cells, "rho")
        if out is None:
            out = torch.zeros(self.total_stations, dtype=torch.float64, device=rho.device)
        _check_t(out, torch.float64, self.total_stations, "g")
        check(lib().ztm_batch_forward(self._h, _ptr(rho), _ptr(out), _stream(stream)))
        return out

    def adjoint(self, r, out=None, stream=None):
        """r float32 [total_stations] -> out float32 [total_cells]."""
        import torch
        _check_t(r, torch.float32, self.total_stations, "r")
        if out is None:
            out = torch.zeros(self.total_cells, dtype=torch.float32, device=r.device)
        _check_t(out, torch.float32, self.total_cells, "out")
        check(lib().ztm_batch_adjoint(self._h, _ptr(r), _ptr(out), _stream(stream)))
        return out

    def solve(self, d, lam, iters=10, x=None, rr=None, use_graph=True, stream=None):
        """CG on (G^T G + lam_s L^T L) x = G^T d per problem, x(0) = 0.
        Returns (x float32 [total_cells], rr float64 [iters + 1, n_problems])."""
        import torch
        _check_t(d, torch.float32, self.total_stations, "d")
        lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64),
                                                   (self.n_problems,)))
        if x is None:
            x = torch.zeros(self.total_cells, dtype=torch.float32, device=d.device)
        _check_t(x, torch.float32, self.total_cells, "x")
        if rr is None:
            rr = torch.zeros(iters + 1, self.n_problems, dtype=torch.float64, device=d.device)
        _check_t(rr, torch.float64, (iters + 1) * self.n_problems, "rr")
        check(lib().ztm_batch_solve(self._h, _ptr(d), lam.ctypes.data_as(ctypes.POINTER(_D)),
                                    iters, _ptr(x), _ptr(rr), int(bool(use_graph)),
                                    _stream(stream)))
        return x, rr

    def set_tuning(self, problem=-1, st_per_lane=2, rows_per_slice=0, cells_per_lane=8,
                   adj_splits=0) -> None:
        check(lib().ztm_batch_set_tuning(self._h, problem, st_per_lane, rows_per_slice,
                                         cells_per_lane, adj_splits))

    def tuning(self, problem: int) -> BatchTuning:
        v = [ctypes.c_int() for _ in range(5)]
        check(lib().ztm_batch_get_tuning(self._h, problem, *[ctypes.byref(a) for a in v]))
        return BatchTuning(*[a.value for a in v])

    def autotune(self, reps: int = 3, stream=None) -> None:
        check(lib().ztm_batch_autotune(self._h, reps, _stream(stream)))


# ------------------------------------------------ dense LA of the GN step
def _f64_cuda(t, shape, name):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous() and tuple(t.shape) == tuple(shape)):
        raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of shape {shape}")


def _gn_work(rows, n, nd, device):
    import torch
    b = lib().ztm_gn_workspace_bytes(rows, n, nd)
    if b < 0:
        raise ValueError(f"invalid sizes rows={rows} n={n} nd={nd}")
    return torch.empty(max(int(b), 1), dtype=torch.uint8, device=device)


def _gn_work_arg(work, rows, n, nd, device):
    """The caller's workspace (a CUDA uint8 tensor of at least
    gn_workspace_bytes(rows, n, nd) bytes) or a new one."""
    import torch
    if work is None:
        return _gn_work(rows, n, nd, device)
    need = lib().ztm_gn_workspace_bytes(rows, n, nd)
    if not (isinstance(work, torch.Tensor) and work.is_cuda and work.dtype == torch.uint8
            and work.is_contiguous() and work.numel() >= need):
        raise ValueError(f"work must be a contiguous uint8 CUDA tensor of >= {need} bytes")
    return work


def gn_workspace_bytes(rows, n, nd=1):
    """Bytes of device workspace the gn_* calls need for J [rows, n] and nd
    damping conditions."""
    b = lib().ztm_gn_workspace_bytes(rows, n, nd)
    if b < 0:
        raise ValueError(f"invalid sizes rows={rows} n={n} nd={nd}")
    return int(b)


def gn_normal(J, r=None, stream=None, out=None, work=None):
    """(J^T J, J^T r) of J float64 [rows, n] and r float64 [rows] on the GPU
    (MTM, zmath.cpp:260-262; T(J)*r, TFM.cpp:160); J^T r is None without r.
    out=(JJ, jtr) and work reuse caller buffers (no allocation)."""
    import torch
    if not (isinstance(J, torch.Tensor) and J.dim() == 2):
        raise ValueError("J must be a 2-D CUDA tensor")
    rows, n = J.shape
    _f64_cuda(J, (rows, n), "J")
    if r is not None:
        _f64_cuda(r, (rows,), "r")
    if out is None:
        JJ = torch.empty(n, n, dtype=torch.float64, device=J.device)
        jtr = torch.empty(n, dtype=torch.float64, device=J.device) if r is not None else None
    else:
        JJ, jtr = out
        _f64_cuda(JJ, (n, n), "JJ")
        if r is not None:
            _f64_cuda(jtr, (n,), "jtr")
    work = _gn_work_arg(work, rows, n, 1, J.device)
    check(lib().ztm_gn_normal(rows, n, _ptr(J), _ptr(r) if r is not None else None, _ptr(JJ),
                              _ptr(jtr) if jtr is not None else None, _ptr(work),
                              _stream(stream)))
    return JJ, jtr


def gn_damped_steps(JJ, jtr, sig, dm, b, stream=None, out=None, work=None):
    """B[k] = b + inv(JJ + dm[k,0] diag(diag(JJ)) + dm[k,1] diag(1/sig^2)) jtr for
    every damping condition dm [nd, 2] (LMA::fit, TFM.cpp:160-161); float64 CUDA
    tensors, B [nd, n]. A system that is not positive definite gives B[k] = b."""
    import torch
    n = JJ.shape[0]
    nd = dm.shape[0]
    _f64_cuda(JJ, (n, n), "JJ")
    _f64_cuda(jtr, (n,), "jtr")
    _f64_cuda(sig, (n,), "sig")
    _f64_cuda(dm, (nd, 2), "dm")
    _f64_cuda(b, (n,), "b")
    if out is None:
        B = torch.empty(nd, n, dtype=torch.float64, device=JJ.device)
    else:
        B = out
        _f64_cuda(B, (nd, n), "out")
    work = _gn_work_arg(work, 0, n, nd, JJ.device)
    check(lib().ztm_gn_damped_steps(n, nd, _ptr(JJ), _ptr(jtr), _ptr(sig), _ptr(dm), _ptr(b),
                                    _ptr(B), _ptr(work), _stream(stream)))
    return B


def gn_inverse_diag(JJ, d, stream=None, out=None, work=None):
    """diag(inv(JJ + I d)) (TFM.cpp:150,173); zeros when not positive definite."""
    import torch
    n = JJ.shape[0]
    _f64_cuda(JJ, (n, n), "JJ")
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=JJ.device)
    else:
        _f64_cuda(out, (n,), "out")
    work = _gn_work_arg(work, 0, n, 1, JJ.device)
    check(lib().ztm_gn_inverse_diag(n, _ptr(JJ), float(d), _ptr(out), _ptr(work),
                                    _stream(stream)))
    return out


def _prec_of(t, n: int, name: str, zoned: bool = False) -> int:
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous() and t.numel() == n):
        raise ValueError(f"{name} must be a contiguous CUDA tensor of {n} values")
    if zoned and t.dtype != torch.float64:
        raise ValueError(f"zoned=True is an FP64 path: {name} must be float64, got {t.dtype}")
    if t.dtype == torch.float32:
        return FP32
    if t.dtype == torch.float64:
        return FP64_ZONED if zoned else FP64
    raise ValueError(f"{name} must be float32 or float64, got {t.dtype}")


def _check_t(t, dtype, n: int, name: str) -> None:
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()
            and t.dtype == dtype and t.numel() == n):
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor of {n} values")


# --------------------------------------------------------------- profiling
def profile_enable(on: bool = True) -> None:
    check(lib().ztm_profile_enable(1 if on else 0))


CALIB_OPS = {"32x32x16_f16": 0, "32x32x16_bf16": 1, "16x16x32_f16": 2}


def calib_mfma(op: str, low_entropy: bool = False, iters: int = 2000, stream=None) -> float:
    """TFLOP/s the chip sustains on one MFMA shape with nothing else in the
    loop (ztm_calib_mfma): op is a key of CALIB_OPS; low_entropy takes
    operands 1 and -0.5 instead of uniform random values."""
    v = ctypes.c_double()
    check(lib().ztm_calib_mfma(CALIB_OPS[op], 1 if low_entropy else 0, iters, ctypes.byref(v),
                               _stream(stream)))
    return v.value


def profile_read(kernel: str) -> tuple[float, int]:
    ms, n = ctypes.c_double(), ctypes.c_int64()
    check(lib().ztm_profile_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def profile_reset() -> None:
    check(lib().ztm_profile_reset())
