"""FP64 CPU oracle for the gravity forward/adjoint path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this package; it is the checker, never the thing measured or shipped. The
product path (ztm_amd, libztm_gfm.so, GFM) does not import it and fails loudly
when its HIP library is missing.

The arithmetic lives in ``gfm_oracle.c`` (see its header for the reference
file:line each function restates and how each is pinned); this module only
loads ``liboracle.so`` (built by ``make -C oracle``) and wraps it for numpy.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# nproc on the GPU host can report far more cores than the share a run gets.
DEFAULT_THREADS = max(1, min(16, os.cpu_count() or 1))


class OGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, f64 = ctypes.c_int, ctypes.c_int64, ctypes.c_double
        gp = ctypes.POINTER(OGrid)
        L.oracle_gf.argtypes = [i32, i32, i32, i32, i32, i32, P, P, P, i32]
        L.oracle_tfm_gravity.argtypes = [i32, i32, i32, i32, i32, i32, P, P, P, i64, P, i32]
        L.oracle_tfm_gravity.restype = i32
        L.oracle_pm_forward.argtypes = [gp, f64, f64, P, i64, P, P, i32]
        L.oracle_pm_adjoint.argtypes = [gp, f64, i64, P, P, P, i32]
        L.oracle_pm_forward_cells.argtypes = [gp, f64, P, P, i64, i64, P, P, i32]
        L.oracle_pm_adjoint_cells.argtypes = [gp, f64, i64, P, P, P, i64, P, i32]
        L.oracle_prism_forward.argtypes = [gp, f64, f64, P, i64, P, P, i32]
        L.oracle_prism_adjoint.argtypes = [gp, f64, i64, P, P, P, i32]
        L.oracle_interp2.argtypes = [P, f64, f64, i32, i32]
        L.oracle_interp2.restype = f64
        L.oracle_interp3.argtypes = [P, f64, f64, f64, i32, i32, i32]
        L.oracle_interp3.restype = f64
        L.oracle_c2g_depth.argtypes = [P, f64, f64, f64]
        L.oracle_c2g_depth.restype = f64
        L.oracle_tfm_drho.argtypes = [i32, i32, i32, i32, i32, P, P, P, P, f64, f64, f64, P, P]
        L.oracle_tfm_drho.restype = i64
        L.oracle_prism_kernel.argtypes = [f64] * 9
        L.oracle_prism_kernel.restype = f64
        L.oracle_mtm.argtypes = [i64, i32, P, P, i32]
        L.oracle_jtr.argtypes = [i64, i32, P, P, P]
        L.oracle_inv.argtypes = [i32, P, P]
        L.oracle_inv.restype = i32
        L.oracle_gn_steps.argtypes = [i32, i32, P, P, P, P, P, P, i32]
        L.oracle_inv_diag.argtypes = [i32, P, f64, P]
        _lib = L
    return _lib


_REF_PATH = os.path.join(_HERE, "_ref", "libzmath_ref.so")
_ref = None


def ref_lib():
    """The reference's own zmath.cpp (MTM, inv) built by ``make -C oracle ref``
    from the reference sources, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        L = ctypes.CDLL(_REF_PATH)
        L.ref_mtm.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_inv.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_inv.restype = ctypes.c_int
        _ref = L
    return _ref


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _grid(g) -> OGrid:
    return OGrid(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz)


def gf(bfr, di, dj, dk, fh, fv, rho, s, g=None, threads=DEFAULT_THREADS) -> np.ndarray:
    """GFM_GPU.cu:16-22 in FP64 on float32 inputs; returns g[di*dj] (accumulated into g)."""
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    s = np.ascontiguousarray(s, dtype=np.float32)
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    if rho.size != n or s.size != di * dj:
        raise ValueError(f"rho needs {n} values and s {di * dj}")
    out = np.zeros(di * dj) if g is None else np.ascontiguousarray(g, dtype=np.float64)
    lib().oracle_gf(bfr, di, dj, dk, fh, fv, _ptr(rho), _ptr(s), _ptr(out), threads)
    return out


def tfm_gravity(bfr, di, dj, dk, fh, fv, drho, surf, nn, G=None, threads=0) -> np.ndarray:
    """TFM.cpp:95-99 CPU gravity (threads=0: one thread per station row)."""
    drho = np.ascontiguousarray(drho, dtype=np.float64)
    surf = np.ascontiguousarray(surf, dtype=np.float64)
    nn = np.ascontiguousarray(nn, dtype=np.int32)
    out = np.zeros(di * dj) if G is None else np.ascontiguousarray(G, dtype=np.float64)
    lib().oracle_tfm_gravity(bfr, di, dj, dk, fh, fv, _ptr(drho), _ptr(surf), _ptr(nn),
                             nn.size, _ptr(out), threads)
    return out


def interp2(f, x, y, di, dj) -> float:
    """zmath.cpp:225-233 (unit spacing) on f[di*dj]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().oracle_interp2(_ptr(f), x, y, di, dj)


def interp3(f, x, y, z, di, dj, dk) -> float:
    """zmath.cpp:236-249 (unit spacing) on f[di*dj*dk]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().oracle_interp3(_ptr(f), x, y, z, di, dj, dk)


def c2g_depth(rc, i, j, k) -> float:
    """Depth (km) of grid point (i, j, k): TFM.cpp:56 + SET.cpp:86-99."""
    rc = np.ascontiguousarray(rc, dtype=np.float64)
    return lib().oracle_c2g_depth(_ptr(rc), i, j, k)


def tfm_drho(di, dj, dk, fh, fv, rc, sea, moho, vel, vm, ro, dr):
    """TFM.cpp:83-89: returns (drho[N] float64, nn int32 ascending)."""
    rc = np.ascontiguousarray(rc, dtype=np.float64)
    sea = np.ascontiguousarray(sea, dtype=np.float64)
    moho = np.ascontiguousarray(moho, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    if sea.size != di * dj or moho.size != di * dj or vel.size != di * dj * dk or rc.size < 7:
        raise ValueError("sea/moho need di*dj values, vel di*dj*dk, rc 7+")
    n = ((di - 1) * fh + 1) * ((dj - 1) * fh + 1) * ((dk - 1) * fv + 1)
    drho = np.empty(n)
    nn = np.empty(n, dtype=np.int32)
    cnt = lib().oracle_tfm_drho(di, dj, dk, fh, fv, _ptr(rc), _ptr(sea), _ptr(moho), _ptr(vel),
                                vm, ro, dr, _ptr(drho), _ptr(nn))
    return drho, nn[:cnt].copy()


def pm_forward(grid, kappa, rho, xyz, threshold=0.0, threads=DEFAULT_THREADS) -> np.ndarray:
    rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    out = np.empty(xyz.shape[0])
    g = _grid(grid)
    lib().oracle_pm_forward(ctypes.byref(g), kappa, threshold, _ptr(rho), xyz.shape[0],
                            _ptr(xyz), _ptr(out), threads)
    return out


def pm_forward_cells(grid, kappa, rho, cells, xyz, threads=DEFAULT_THREADS) -> np.ndarray:
    """Forward restricted to the listed cells (bounded CPU-baseline sample)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    out = np.empty(xyz.shape[0])
    g = _grid(grid)
    lib().oracle_pm_forward_cells(ctypes.byref(g), kappa, _ptr(rho), _ptr(cells), cells.size,
                                  xyz.shape[0], _ptr(xyz), _ptr(out), threads)
    return out


def pm_adjoint(grid, kappa, xyz, r, threads=DEFAULT_THREADS) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    out = np.empty(grid.nx * grid.ny * grid.nz)
    g = _grid(grid)
    lib().oracle_pm_adjoint(ctypes.byref(g), kappa, xyz.shape[0], _ptr(xyz), _ptr(r), _ptr(out),
                            threads)
    return out


def pm_adjoint_cells(grid, kappa, xyz, r, cells, threads=DEFAULT_THREADS) -> np.ndarray:
    """Adjoint restricted to the listed cells (bounded CPU-baseline sample)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    out = np.empty(cells.size)
    g = _grid(grid)
    lib().oracle_pm_adjoint_cells(ctypes.byref(g), kappa, xyz.shape[0], _ptr(xyz), _ptr(r),
                                  _ptr(cells), cells.size, _ptr(out), threads)
    return out


def prism_forward(grid, kappa, rho, xyz, threshold=0.0, threads=DEFAULT_THREADS) -> np.ndarray:
    rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    out = np.empty(xyz.shape[0])
    g = _grid(grid)
    lib().oracle_prism_forward(ctypes.byref(g), kappa, threshold, _ptr(rho), xyz.shape[0],
                               _ptr(xyz), _ptr(out), threads)
    return out


def prism_adjoint(grid, kappa, xyz, r, threads=DEFAULT_THREADS) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    out = np.empty(grid.nx * grid.ny * grid.nz)
    g = _grid(grid)
    lib().oracle_prism_adjoint(ctypes.byref(g), kappa, xyz.shape[0], _ptr(xyz), _ptr(r),
                               _ptr(out), threads)
    return out


def prism_kernel(xs, ys, zs, x1, x2, y1, y2, z1, z2) -> float:
    """Integral of (zs - z)/r^3 over [x1,x2]x[y1,y2]x[z1,z2]."""
    return lib().oracle_prism_kernel(xs, ys, zs, x1, x2, y1, y2, z1, z2)


# ------------------------------------------------ dense LA of the GN step
def mtm(J, threads=DEFAULT_THREADS) -> np.ndarray:
    """J^T J (zmath.cpp:260-262) of J [rows, cols]."""
    J = np.ascontiguousarray(J, dtype=np.float64)
    out = np.empty((J.shape[1], J.shape[1]))
    lib().oracle_mtm(J.shape[0], J.shape[1], _ptr(J), _ptr(out), threads)
    return out


def jtr(J, r) -> np.ndarray:
    """J^T r (TFM.cpp:160)."""
    J = np.ascontiguousarray(J, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(J.shape[1])
    lib().oracle_jtr(J.shape[0], J.shape[1], _ptr(J), _ptr(r), _ptr(out))
    return out


def inv(a):
    """(ok, inverse) by zmath.cpp:129-140's Gauss-Jordan; ok False and a zero
    matrix on an exact zero pivot."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    ok = lib().oracle_inv(a.shape[0], _ptr(a), _ptr(out))
    return bool(ok), out


def gn_steps(JJ, jtr_, sig, dm, b, threads=DEFAULT_THREADS) -> np.ndarray:
    """B[k] = b + inv(JJ + dm[k,0] diag(diag(JJ)) + dm[k,1] diag(1/sig^2)) jtr
    (TFM.cpp:159-161)."""
    JJ = np.ascontiguousarray(JJ, dtype=np.float64)
    n = JJ.shape[0]
    jtr_ = np.ascontiguousarray(jtr_, dtype=np.float64)
    sig = np.ascontiguousarray(sig, dtype=np.float64)
    dm = np.ascontiguousarray(dm, dtype=np.float64).reshape(-1, 2)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty((dm.shape[0], n))
    lib().oracle_gn_steps(n, dm.shape[0], _ptr(JJ), _ptr(jtr_), _ptr(sig), _ptr(dm), _ptr(b),
                          _ptr(out), threads)
    return out


def inv_diag(JJ, d) -> np.ndarray:
    """diag(inv(JJ + I d)) (TFM.cpp:150,173); zeros when singular."""
    JJ = np.ascontiguousarray(JJ, dtype=np.float64)
    out = np.empty(JJ.shape[0])
    lib().oracle_inv_diag(JJ.shape[0], _ptr(JJ), float(d), _ptr(out))
    return out
