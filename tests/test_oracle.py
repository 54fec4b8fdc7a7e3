"""Pins for the FP64 oracle (oracle/gfm_oracle.c), CPU only.

Each check is chosen so that a plausible mistake fails it: a wrong constant or
unit (closed forms), a sign flip (mass below -> g > 0), a transposed or
mis-decoded raster index (mirror symmetry, brute force with an independently
built cell list, dot-product identity between two independent loop nests), a
dropped term (solid-angle and slab limits), a wrong threshold (the .001 skip).
"""
import math

import numpy as np
import pytest

from ztm_amd import workloads as W

G_SI = 6.674e-11


def grid1(nx, ny, nz, x0=0.0, y0=0.0, z0=0.0, dx=1.0, dy=1.0, dz=1.0):
    return W.Grid(nx, ny, nz, x0, y0, z0, dx, dy, dz)


# ------------------------------------------------------------------ units
def test_reference_constant_is_G_times_units():
    # 66.7 = G[SI] * 1e3 (g/cc -> kg/m^3) * 1e4 (h = 10 km, TFM.cpp:190) / 1e-5 (mGal)
    assert abs(6.67e-11 * 1e3 * 1e4 / 1e-5 - 66.7) < 1e-9
    assert abs(W.KAPPA_SI - 1e5 * G_SI) < 1e-20


def test_single_cell_is_newton(orc):
    """1 km^3 of +1000 kg/m^3 10 km below a station: g = G M / r^2 = 0.06674 mGal."""
    g = grid1(1, 1, 1, 0.0, 0.0, -10000.0, 1000.0, 1000.0, 1000.0)
    out = orc.pm_forward(g, W.KAPPA_SI, np.array([1000.0]), np.array([[0.0, 0.0, 0.0]]))
    newton = G_SI * 1000.0 * 1e9 / 1e8 / 1e-5
    assert abs(out[0] - newton) < 1e-12 * newton
    # sign: the same mass above the station pulls up (g_z < 0)
    up = orc.pm_forward(g, W.KAPPA_SI, np.array([1000.0]), np.array([[0.0, 0.0, -20000.0]]))
    assert abs(up[0] + newton) < 1e-12 * newton


def test_single_cell_off_axis(orc):
    g = grid1(1, 1, 1, 300.0, -400.0, -1200.0, 10.0, 20.0, 5.0)
    st = np.array([[0.0, 0.0, 0.0]])
    out = orc.pm_forward(g, 2.0, np.array([3.0]), st)
    r = math.sqrt(300.0 ** 2 + 400.0 ** 2 + 1200.0 ** 2)
    assert abs(out[0] - 2.0 * 3.0 * 1000.0 * 1200.0 / r ** 3) < 1e-15


def test_gf_single_cell_reference_units(orc):
    """GFM_GPU.cu:22 in index units: one cell straight below an interior station."""
    di = dj = 7
    dk, fh, fv, bfr = 4, 1, 2, 2
    kk = (dk - 1) * fv + 1
    rho = np.zeros(di * dj * kk, np.float32)
    i, j, k = 3, 3, 1                          # cell at (3, 3, k/fv = 0.5)
    rho[i + j * di + k * di * dj] = 0.25
    s = np.full(di * dj, 2.5, np.float32)       # stations at k0 = 2.5
    g = orc.gf(bfr, di, dj, dk, fh, fv, rho, s)
    m = 3 + 3 * di
    assert abs(g[m] - 66.7 / (fh * fh * fv) * 0.25 / 2.0 ** 2) < 1e-12
    # neighbour station (4,3): r^2 = 1 + 4
    assert abs(g[4 + 3 * di] - 66.7 / 2 * 0.25 * 2.0 / 5.0 ** 1.5) < 1e-12
    # exterior stations (within bfr of the edge) are not computed
    assert g[0] == 0.0 and g[1 + 1 * di] == 0.0


def test_gf_threshold_skips_small_cells(orc):
    di = dj = 9
    dk, fh, fv, bfr = 3, 1, 2, 2
    n = di * dj * ((dk - 1) * fv + 1)
    rho = W.uniform(1, 1, n, -0.01, 0.01).astype(np.float32)
    rho[::7] = np.float32(0.001)               # exactly at the threshold
    rho[::11] = np.float32(-0.0009)
    s = W.uniform(1, 2, di * dj, 3.0, 6.0).astype(np.float32)
    g = orc.gf(bfr, di, dj, dk, fh, fv, rho, s)
    kept = np.where(np.abs(rho) > np.float32(0.001), rho, 0).astype(np.float32)
    assert np.array_equal(g, orc.gf(bfr, di, dj, dk, fh, fv, kept, s))
    assert not np.array_equal(g, orc.gf(bfr, di, dj, dk, fh, fv, rho * 2, s))


def test_gf_equals_tfm_cpu_lambda(orc):
    """GFM_GPU.cu:16-22 and TFM.cpp:95-99 restatements agree on shared inputs."""
    di, dj, dk, fh, fv, bfr = 15, 13, 5, 1, 2, 3
    n = di * dj * ((dk - 1) * fv + 1)
    rho = W.uniform(2, 1, n, -0.05, 0.05).astype(np.float32)
    s = W.uniform(2, 2, di * dj, 6.0, 9.0).astype(np.float32)
    nn = np.nonzero(np.abs(rho) > np.float32(0.001))[0].astype(np.int32)
    a = orc.gf(bfr, di, dj, dk, fh, fv, rho, s)
    b = orc.tfm_gravity(bfr, di, dj, dk, fh, fv, rho.astype(np.float64), s.astype(np.float64), nn)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    c = orc.tfm_gravity(bfr, di, dj, dk, fh, fv, rho.astype(np.float64), s.astype(np.float64), nn,
                        threads=3)
    assert np.array_equal(b, c)


# ------------------------------------------------------- geometry checks
def test_mirror_symmetry(orc):
    g = grid1(6, 5, 4, 0.5, 0.5, -0.5, 1.0, 1.0, -1.0)
    rho = W.uniform(3, 1, g.n_cells, -1, 1).reshape(4, 5, 6)
    st = np.array([[1.3, 2.2, 0.4], [4.9, 0.7, 1.5]])
    a = orc.pm_forward(g, 1.0, rho.ravel(), st)
    # mirror x: cell i -> 5 - i, station x -> 6 - x
    b = orc.pm_forward(g, 1.0, rho[:, :, ::-1].ravel(), np.c_[6.0 - st[:, 0], st[:, 1:]])
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))
    # mirror y
    c = orc.pm_forward(g, 1.0, rho[:, ::-1, :].ravel(), np.c_[st[:, 0], 5.0 - st[:, 1], st[:, 2]])
    assert np.max(np.abs(a - c)) < 1e-13 * np.max(np.abs(a))
    # a transposed raster would break this: swap x/y on a non-square grid fails
    t = orc.pm_forward(g, 1.0, rho.transpose(0, 2, 1).ravel(), st)
    assert np.max(np.abs(a - t)) > 1e-3 * np.max(np.abs(a))


def test_brute_force_tiny(orc):
    g = grid1(3, 2, 2, 10.0, -5.0, 2.0, 2.0, 3.0, -1.5)
    rho = W.uniform(4, 1, g.n_cells, -2, 2)
    st = np.array([[11.0, -4.0, 5.0], [14.0, 1.0, -1.0], [9.0, -6.0, 3.0]])
    want = np.zeros(3)
    V = 2.0 * 3.0 * 1.5
    n = 0
    for k in range(2):
        for j in range(2):
            for i in range(3):
                c = np.array([10.0 + 2.0 * i, -5.0 + 3.0 * j, 2.0 - 1.5 * k])
                for m in range(3):
                    d = st[m] - c
                    want[m] += rho[n] * V * d[2] / np.linalg.norm(d) ** 3
                n += 1
    got = orc.pm_forward(g, 1.0, rho, st)
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


def _square_solid_angle(a, d):
    """Solid angle of a 2a x 2a square seen from height d above its centre."""
    return 4.0 * math.atan(a * a / (d * math.sqrt(2.0 * a * a + d * d)))


def test_thin_layer_solid_angle(orc):
    """A thin uniform layer: g = G rho t Omega, Omega = solid angle of the square."""
    h, n, d = 100.0, 160, 3000.0
    g = W.Grid(n, n, 1, -0.5 * (n - 1) * h, -0.5 * (n - 1) * h, -d, h, h, h)
    rho = np.full(g.n_cells, 500.0)
    out = orc.pm_forward(g, W.KAPPA_SI, rho, np.array([[0.0, 0.0, 0.0]]))
    want = W.KAPPA_SI * 500.0 * h * _square_solid_angle(0.5 * n * h, d)
    assert abs(out[0] / want - 1.0) < 1e-3


# ------------------------------------------------------------- adjoint
def test_adjoint_dot_product(orc):
    w = W.adjoint_small()
    r = w.extra["residual"]
    g = orc.pm_forward(w.grid, w.kappa, w.rho, w.stations)
    a = orc.pm_adjoint(w.grid, w.kappa, w.stations, r)
    lhs, rhs = g @ r, w.rho @ a
    assert abs(lhs - rhs) < 1e-12 * (np.abs(g) @ np.abs(r))


def test_adjoint_is_transpose_on_anisotropic_grid(orc):
    g = W.Grid(7, 4, 3, 1.0, 2.0, -0.5, 1.5, -0.7, -2.0)
    st = np.stack([W.uniform(6, a, 11, -3, 9) for a in range(3)], axis=1)
    rho = W.uniform(6, 7, g.n_cells, -1, 1)
    r = W.normal(6, 8, 11, 1.0)
    lhs = orc.pm_forward(g, 1.3, rho, st) @ r
    rhs = rho @ orc.pm_adjoint(g, 1.3, st, r)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    cells = np.array([0, 5, 40, g.n_cells - 1])
    assert np.allclose(orc.pm_adjoint_cells(g, 1.3, st, r, cells),
                       orc.pm_adjoint(g, 1.3, st, r)[cells], rtol=0, atol=1e-14)


def test_forward_over_listed_cells_is_the_masked_forward(orc):
    """oracle_pm_forward_cells against the pinned full forward: the sum over a
    cell list equals the forward of rho masked to that list (a mis-decoded
    cell or rho read at the list position instead of the cell fails), the full
    list equals the forward itself, and the sums over the two halves of a
    shuffled partition add up to the whole."""
    g = W.Grid(7, 4, 3, 1.0, 2.0, -0.5, 1.5, -0.7, -2.0)
    st = np.stack([W.uniform(16, a, 11, -3, 9) for a in range(3)], axis=1)
    rho = W.uniform(16, 7, g.n_cells, -1, 1)
    full = orc.pm_forward(g, 1.3, rho, st)
    scale = np.abs(full).max()
    allc = np.arange(g.n_cells)
    assert np.abs(orc.pm_forward_cells(g, 1.3, rho, allc, st) - full).max() <= 1e-14 * scale
    cells = np.array([3, 17, 40, 41, 66, g.n_cells - 1])
    masked = np.zeros_like(rho)
    masked[cells] = rho[cells]
    sub = orc.pm_forward_cells(g, 1.3, rho, cells, st)
    assert np.abs(sub - orc.pm_forward(g, 1.3, masked, st)).max() <= 1e-14 * scale
    assert np.abs(sub).max() > 1e-3 * scale
    perm = (W.uniform(16, 9, g.n_cells) * 1e6).argsort()
    a, b = perm[: g.n_cells // 3], perm[g.n_cells // 3:]
    parts = orc.pm_forward_cells(g, 1.3, rho, a, st) + orc.pm_forward_cells(g, 1.3, rho, b, st)
    assert np.abs(parts - full).max() <= 1e-13 * scale


def grad_laplacian_term(rho, shape):
    """Gradient of 0.5 * ||grad rho||^2 with forward differences on each axis."""
    f = rho.reshape(shape)
    out = np.zeros_like(f)
    for ax in range(3):
        d = np.diff(f, axis=ax)
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(0, -1)
        sl_hi[ax] = slice(1, None)
        out[tuple(sl_lo)] -= d
        out[tuple(sl_hi)] += d
    return out.ravel()


def reg_local(rho, shape, n):
    """The terms of 0.5 * ||grad rho||^2 (forward differences) that involve cell n;
    the others do not change when only rho[n] does."""
    f = rho.reshape(shape)
    k, j, i = np.unravel_index(n, shape)
    s = 0.0
    for ax, idx in enumerate((k, j, i)):
        for o in (-1, 1):
            q = [k, j, i]
            q[ax] = idx + o
            if 0 <= q[ax] < shape[ax]:
                s += 0.5 * (f[k, j, i] - f[tuple(q)]) ** 2
    return s


def test_misfit_gradient_central_differences(orc):
    """grad of 0.5||G rho - d||^2 + 0.5 lam ||grad rho||^2 = G^T r + lam L^T L rho,
    against central differences (h = 1e-3 kg/m^3) at 64 seeded cells."""
    w = W.adjoint_small()
    lam = w.extra["lambda"]
    shape = (w.grid.nz, w.grid.ny, w.grid.nx)
    g0 = orc.pm_forward(w.grid, w.kappa, w.rho, w.stations)
    d = g0 - w.extra["residual"]                      # residual at rho0 is r
    res = g0 - d
    grad = orc.pm_adjoint(w.grid, w.kappa, w.stations, res) + lam * grad_laplacian_term(w.rho, shape)
    cells = np.unique((W.uniform(99, 1, 64) * w.grid.n_cells).astype(int))
    h = 1e-3
    for n in cells:
        e = np.zeros(w.grid.n_cells)
        e[n] = 1.0
        col = orc.pm_forward(w.grid, w.kappa, e, w.stations)     # G e_n, an independent forward
        fp = 0.5 * np.sum((g0 + h * col - d) ** 2) + lam * reg_local(w.rho + h * e, shape, n)
        fm = 0.5 * np.sum((g0 - h * col - d) ** 2) + lam * reg_local(w.rho - h * e, shape, n)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - grad[n]) <= 1e-6 * max(abs(grad[n]), 1e-3), (n, fd, grad[n])


# ---------------------------------------------------------------- prism
def _gl_box(xs, ys, zs, box, n=48):
    x1, x2, y1, y2, z1, z2 = box
    t, w = np.polynomial.legendre.leggauss(n)
    X = (x1 + x2) / 2 + (x2 - x1) / 2 * t
    Y = (y1 + y2) / 2 + (y2 - y1) / 2 * t
    Z = (z1 + z2) / 2 + (z2 - z1) / 2 * t
    WX, WY, WZ = w * (x2 - x1) / 2, w * (y2 - y1) / 2, w * (z2 - z1) / 2
    x, y, z = np.meshgrid(X, Y, Z, indexing="ij")
    W3 = WX[:, None, None] * WY[None, :, None] * WZ[None, None, :]
    r = np.sqrt((xs - x) ** 2 + (ys - y) ** 2 + (zs - z) ** 2)
    return float(np.sum(W3 * (zs - z) / r ** 3))


@pytest.mark.parametrize("st", [(0.0, 0.0, 1.0), (0.3, 0.2, 2.0), (3.0, 1.0, 0.5),
                                (0.0, 0.0, -3.0), (0.7, -0.2, -0.4), (-5.0, 4.0, -2.5)])
def test_prism_matches_quadrature(orc, st):
    box = (-0.5, 0.5, -0.4, 0.4, -1.0, 0.0)
    want = _gl_box(*st, box)
    got = orc.prism_kernel(*st, *box)
    assert abs(got - want) < 1e-9 * max(abs(want), 1e-3)


def test_prism_far_field_is_point_mass(orc):
    box = (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5)       # unit cube at the origin
    errs = []
    for r in (4.0, 8.0, 16.0):
        st = (0.3 * r, 0.2 * r, math.sqrt(1 - 0.13) * r)
        pm = st[2] / r ** 3
        errs.append(abs(orc.prism_kernel(*st, *box) / pm - 1.0))
    assert errs[-1] < 1e-4
    # a cube has no quadrupole moment: the error falls off at least as (a/r)^4
    assert errs[1] / errs[0] < 1 / 12 and errs[2] / errs[1] < 1 / 12


def test_prism_subdivision_additive(orc):
    st = (0.37, -0.21, 0.8)
    box = (-0.5, 0.5, -0.5, 0.5, -1.0, 0.0)
    whole = orc.prism_kernel(*st, *box)
    parts = 0.0
    for xa, xb in ((-0.5, 0.0), (0.0, 0.5)):
        for ya, yb in ((-0.5, 0.0), (0.0, 0.5)):
            for za, zb in ((-1.0, -0.5), (-0.5, 0.0)):
                parts += orc.prism_kernel(*st, xa, xb, ya, yb, za, zb)
    assert abs(whole - parts) < 1e-12 * abs(whole)


def test_prism_slab_limit(orc):
    """A wide plate of thickness t: g = G rho int Omega(z) dz -> 2 pi G rho t."""
    a, t, d = 2.0e5, 100.0, 1000.0
    got = orc.prism_kernel(0.0, 0.0, 0.0, -a, a, -a, a, -d - t, -d)
    zz, ww = np.polynomial.legendre.leggauss(20)
    z = d + t / 2 + t / 2 * zz
    want = sum(wi * t / 2 * _square_solid_angle(a, zi) for zi, wi in zip(z, ww))
    assert abs(got - want) < 1e-9 * want
    assert abs(got / (2 * math.pi * t) - 1.0) < 1e-2


def test_prism_grid_dot_product(orc):
    g = W.Grid(5, 4, 3, 0.5, 0.5, -0.5, 1.0, 1.0, -1.0)
    st = np.stack([W.uniform(8, a, 9, -1, 6) for a in range(2)] + [W.uniform(8, 2, 9, 0, 2)],
                  axis=1)
    rho = W.uniform(8, 5, g.n_cells, -1, 1)
    r = W.normal(8, 6, 9, 1.0)
    lhs = orc.prism_forward(g, 1.0, rho, st) @ r
    rhs = rho @ orc.prism_adjoint(g, 1.0, st, r)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
