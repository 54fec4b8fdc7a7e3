"""Seeded synthetic gravity workloads (grids, stations, density models).

Every random draw comes from a counter-based SplitMix64 stream, so any value
can be regenerated from (seed, stream, index) alone and the host and any test
see identical inputs. Seeds are 20190301 + workload id (pm_cpu_small 1,
pm_blobs_1m 2, adjoint_small 3, ensemble_256 4, uneven_48 5, tfm_fm 6,
gn_step 7).

Coordinates are metres with z up; cells of edge h fill x in [0, nx h],
y in [0, ny h], z in [-nz h, 0] (top at z = 0); density contrast in kg/m^3;
the gravity output is in mGal (kappa = 1e5 * G).

    pm_cpu_small   32x32x16 cells of 1 km, 1,024 stations at z=+10 m, FP64
                   U[-300,300] densities (the CPU-runnable case)
    pm_blobs_1m    128x128x64 cells of 500 m (1,048,576), 8 Gaussian blobs
                   (+-400 kg/m^3, sigma 2-8 km) + N(0,20) noise; 16,384 stations
                   on a 128x128 grid, heights 0-2,000 m from 5 Gaussian hills
    adjoint_small  pm_cpu_small geometry, residual N(0, 0.5) mGal, lambda 1e-3
    ensemble_256   96x96x48 cells of 1 km, 256 models of U[-300,300]
                   smoothed by a clamped 3x3x3 box filter; 9,216 stations, z=+10
    tfm_fm         the reference's own gravity FM call (TFM.cpp:83-94) at its
                   grid (61 x 61 x 33 coarse, h = 10 km, fh = 1, fv = 2, bfr 10,
                   reference point (-15.25, -72.5)): sea-level k from the ENU
                   frame, seeded topography, Airy Moho + perturbation,
                   perturbed mantle velocity; drho rendered from them
    uneven_48      48 independent problems: cells log-U[2e4, 2e6], stations
                   log-U[1e3, 4e4] scattered over the footprint at z=+10 m,
                   cell edge log-U[250, 2000] m; 3 buried +250 kg/m^3 blocks
                   each; data noise N(0, 0.1 mGal)
    gn_step        one damped Gauss-Newton step of LMA::fit (TFM.cpp:157-161):
                   Jacobian 32,768 data x 1,024 parameters (column scales
                   e^U[-3,3], 30 % zeros), residual N(0,1), a-priori sigma
                   U[0.05, 4], the reference's 12 damping conditions
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED0 = 20190301
G_SI = 6.674e-11
KAPPA_SI = 1e5 * G_SI          # mGal per (kg/m^3 * m^3 / m^2)

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def _splitmix(x: np.ndarray) -> np.ndarray:
    z = x + _M1
    z = (z ^ (z >> np.uint64(30))) * _M2
    z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """n uniforms in [lo, hi) from SplitMix64(seed, stream, index)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    idx = np.arange(n, dtype=np.uint64) * _M1
    with np.errstate(over="ignore"):
        bits = _splitmix(idx + base)
    u = (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, stream: int, n: int, sigma: float = 1.0) -> np.ndarray:
    """n normals N(0, sigma^2) by Box-Muller on two uniform streams."""
    u1 = uniform(seed, 2 * stream + 1000, n)
    u2 = uniform(seed, 2 * stream + 1001, n)
    return sigma * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    x0: float
    y0: float
    z0: float
    dx: float
    dy: float
    dz: float

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def centres(self) -> np.ndarray:
        """[N,3] cell centres in raster order n = i + nx*(j + ny*k)."""
        k, j, i = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx),
                              indexing="ij")
        return np.stack([self.x0 + i.ravel() * self.dx, self.y0 + j.ravel() * self.dy,
                         self.z0 + k.ravel() * self.dz], axis=1)


def cube_grid(nx: int, ny: int, nz: int, h: float) -> Grid:
    """Cubic cells of edge h, top face at z = 0, k counting downwards."""
    return Grid(nx, ny, nz, 0.5 * h, 0.5 * h, -0.5 * h, h, h, -h)


@dataclass
class Workload:
    name: str
    grid: Grid
    stations: np.ndarray            # [M,3] float64
    rho: np.ndarray                 # [N] or [B,N] float64
    kappa: float = KAPPA_SI
    extra: dict = field(default_factory=dict)

    @property
    def n_stations(self) -> int:
        return int(self.stations.shape[0])

    @property
    def pairs(self) -> int:
        return self.n_stations * self.grid.n_cells


def _station_grid(g: Grid, z: np.ndarray | float) -> np.ndarray:
    j, i = np.meshgrid(np.arange(g.ny), np.arange(g.nx), indexing="ij")
    x = g.x0 + i.ravel() * g.dx
    y = g.y0 + j.ravel() * g.dy
    zz = np.broadcast_to(np.asarray(z, dtype=np.float64), x.shape).astype(np.float64)
    return np.stack([x, y, zz], axis=1)


def pm_cpu_small(seed: int = SEED0 + 1) -> Workload:
    g = cube_grid(32, 32, 16, 1000.0)
    rho = uniform(seed, 1, g.n_cells, -300.0, 300.0)
    return Workload("pm_cpu_small", g, _station_grid(g, 10.0), rho)


def _hills(seed: int, x: np.ndarray, y: np.ndarray, extent: float) -> np.ndarray:
    u = uniform(seed, 20, 5 * 4).reshape(5, 4)
    z = np.zeros_like(x)
    for a, s, cx, cy in u:
        amp = 500.0 + 1000.0 * a
        sig = extent * (0.05 + 0.15 * s)
        z += amp * np.exp(-((x - cx * extent) ** 2 + (y - cy * extent) ** 2) / (2 * sig * sig))
    return np.clip(z, 0.0, 2000.0)


def pm_blobs_1m(seed: int = SEED0 + 2, n: int = 128, nz: int = 64, h: float = 500.0) -> Workload:
    g = cube_grid(n, n, nz, h)
    ex, ez = n * h, nz * h
    u = uniform(seed, 10, 8 * 6).reshape(8, 6)
    c = g.centres()
    rho = np.zeros(g.n_cells)
    for sgn, sig, cx, cy, cz, _ in u:
        amp = 400.0 if sgn < 0.5 else -400.0
        s = 2000.0 + 6000.0 * sig
        d2 = (c[:, 0] - cx * ex) ** 2 + (c[:, 1] - cy * ex) ** 2 + (c[:, 2] + cz * ez) ** 2
        rho += amp * np.exp(-d2 / (2.0 * s * s))
    rho += normal(seed, 11, g.n_cells, 20.0)
    st = _station_grid(g, 0.0)
    st[:, 2] = _hills(seed, st[:, 0], st[:, 1], ex)
    return Workload("pm_blobs_1m", g, st, rho)


def adjoint_small(seed: int = SEED0 + 3) -> Workload:
    w = pm_cpu_small(SEED0 + 1)
    r = normal(seed, 30, w.n_stations, 0.5)
    return Workload("adjoint_small", w.grid, w.stations, w.rho,
                    extra={"residual": r, "lambda": 1e-3})


def _box3(field3: np.ndarray) -> np.ndarray:
    """3x3x3 mean with clamped (edge-replicated) borders over the last 3 axes."""
    out = field3
    for ax in (-1, -2, -3):
        p = np.concatenate([np.take(out, [0], axis=ax), out, np.take(out, [-1], axis=ax)],
                           axis=ax)
        n = out.shape[ax]
        sl = [slice(None)] * out.ndim
        acc = 0.0
        for o in range(3):
            sl[ax] = slice(o, o + n)
            acc = acc + p[tuple(sl)]
        out = acc / 3.0
    return out


def ensemble_256(seed: int = SEED0 + 4, models: int = 256, n: int = 96, nz: int = 48,
                 h: float = 1000.0, dtype=np.float32) -> Workload:
    g = cube_grid(n, n, nz, h)
    rho = np.empty((models, g.n_cells), dtype=dtype)
    for b in range(models):
        f = uniform(seed, 100 + b, g.n_cells, -300.0, 300.0).reshape(nz, n, n)
        rho[b] = _box3(f).ravel()
    return Workload("ensemble_256", g, _station_grid(g, 10.0), rho)


def tfm_shaped(seed: int = 7, di: int = 61, dj: int = 61, dk: int = 33, fh: int = 1,
               fv: int = 2):
    """Inputs of the reference's own gravity call (TFM.cpp:83-94, GFM_GPU.cu:24):
    float32 drho[ii*jj*kk] in g/cc on the fine raster, a third of it at or
    below the .001 g/cc skip threshold (two cells exactly at +-.001), and
    float32 station heights s[di*dj] in k-index units around the grid top.
    Defaults are TFM's grid (di=dj=61, dk=33, fh=1, fv=2; TFM.cpp:190)."""
    ii, jj, kk = (di - 1) * fh + 1, (dj - 1) * fh + 1, (dk - 1) * fv + 1
    n = ii * jj * kk
    rho = uniform(seed, 1, n, -0.3, 0.3)
    small = uniform(seed, 2, n) < 0.33
    rho[small] = uniform(seed, 3, int(small.sum()), -0.001, 0.001)
    rho = rho.astype(np.float32)
    rho[5] = np.float32(0.001)
    rho[6] = np.float32(-0.001)
    s = uniform(seed, 4, di * dj, kk / fv - 4.0, kk / fv + 1.5).astype(np.float32)
    return rho, s


@dataclass
class Subproblem:
    grid: Grid
    stations: np.ndarray            # [M,3] float64
    rho: np.ndarray                 # [N] float64, the true model (blocks)
    noise: np.ndarray               # [M] float64, mGal
    kappa: float = KAPPA_SI

    @property
    def lam(self) -> float:
        """Smoothness weight: 0.05 * 2 pi M (kappa h / nx)^2, a fixed fraction of
        the squared column norm of G for a top cell under M stations spread
        over the nx-cell footprint (recorded choice)."""
        m = int(self.stations.shape[0])
        return 0.05 * 2.0 * np.pi * m * (self.kappa * self.grid.dx / self.grid.nx) ** 2

    @property
    def pairs(self) -> int:
        return int(self.stations.shape[0]) * self.grid.n_cells


def recovery_stats(x: np.ndarray, p: Subproblem) -> dict:
    """How well a recovered model x [N] (cell order of p.grid) finds p's
    buried blocks, from the column masses c = sum over depth (gravity fixes
    the horizontal distribution of mass; its depth is left to the prior):
    corr, the correlation of recovered and true column masses; mass, the
    recovered total excess mass over the true one; contrast, the mean
    recovered column mass over the blocks' footprints against that outside
    them (mean |c| outside, so that a model of alternating sign is not
    rewarded); offset, the distance between the recovered and the true
    mass centroids (of c clipped at 0) in units of the grid's width."""
    g = p.grid
    c = x.reshape(g.nz, g.ny, g.nx).sum(axis=0)
    t = p.rho.reshape(g.nz, g.ny, g.nx).sum(axis=0)
    inside = t > 0
    cin = float(c[inside].mean())
    cout = float(np.abs(c[~inside]).mean()) if (~inside).any() else 0.0
    jj, ii = np.meshgrid(np.arange(g.ny), np.arange(g.nx), indexing="ij")

    def centroid(w):
        w = np.clip(w, 0.0, None)
        return np.array([(w * ii).sum(), (w * jj).sum()]) / w.sum()

    return {"corr": float(np.corrcoef(c.ravel(), t.ravel())[0, 1]),
            "mass": float(c.sum() / t.sum()),
            "contrast": cin / cout if cout > 0 else float("inf"),
            "offset": float(np.linalg.norm(centroid(c) - centroid(t)) / g.nx)}


def _logu(seed: int, stream: int, lo: float, hi: float) -> float:
    return float(np.exp(uniform(seed, stream, 1, np.log(lo), np.log(hi))[0]))


def uneven_48(seed: int = SEED0 + 5, n_problems: int = 48, cells=(2e4, 2e6),
              stations=(1e3, 4e4), edge=(250.0, 2000.0), noise: float = 0.1):
    """Problems of uneven size for the batched solver. Problem s draws from
    streams 100*s + (0..19): cell count N, station count M and edge h
    log-uniform in the given ranges; the grid is nx = ny = 2 nz cubes with
    nx*ny*nz ~ N; stations uniform over the footprint at z = +10 m; the true
    model is three boxes of +250 kg/m^3 (each 10-30 % of every axis, tops at
    least one cell below the surface) on zero background."""
    out = []
    for s in range(n_problems):
        b = 100 * s
        n = _logu(seed, b, *cells)
        m = int(round(_logu(seed, b + 1, *stations)))
        h = _logu(seed, b + 2, *edge)
        nz = max(4, int(round((n / 4.0) ** (1.0 / 3.0))))
        nx = max(4, int(round(np.sqrt(n / nz))))
        g = cube_grid(nx, nx, nz, h)
        xy = uniform(seed, b + 3, 2 * m, 0.0, nx * h).reshape(m, 2)
        st = np.concatenate([xy, np.full((m, 1), 10.0)], axis=1)
        rho = np.zeros((nz, nx, nx))
        u = uniform(seed, b + 4, 3 * 6).reshape(3, 6)
        for fx, fy, fz, px, py, pz in u:
            sx = max(1, int(round(nx * (0.1 + 0.2 * fx))))
            sy = max(1, int(round(nx * (0.1 + 0.2 * fy))))
            sz = max(1, int(round(nz * (0.1 + 0.2 * fz))))
            i0 = int(px * (nx - sx + 1))
            j0 = int(py * (nx - sy + 1))
            k0 = 1 + int(pz * max(1, nz - sz))
            rho[k0:k0 + sz, j0:j0 + sy, i0:i0 + sx] = 250.0
        out.append(Subproblem(g, st, rho.ravel(), normal(seed, b + 5, m, noise)))
    return out


_RAD, _R_EARTH = 57.295, 6371.0     # SET.cpp:67,87


def _geo2enu(ref, lat, lon, d):
    """SET.cpp:66-82 (vectorised): geodetic (deg, deg, km depth) -> ENU km."""
    lat_r, lon_r, d_r = ref
    xi = (_R_EARTH - d) * np.cos(lon / _RAD) * np.cos(lat / _RAD)
    yi = (_R_EARTH - d) * np.sin(lon / _RAD) * np.cos(lat / _RAD)
    zi = (_R_EARTH - d) * np.sin(lat / _RAD)
    xr = (_R_EARTH - d_r) * np.cos(lon_r / _RAD) * np.cos(lat_r / _RAD)
    yr = (_R_EARTH - d_r) * np.sin(lon_r / _RAD) * np.cos(lat_r / _RAD)
    zr = (_R_EARTH - d_r) * np.sin(lat_r / _RAD)
    ph, th = lat_r / _RAD, lon_r / _RAD
    v = (xi - xr, yi - yr, zi - zr)
    x = -np.sin(th) * v[0] + np.cos(th) * v[1]
    y = -np.sin(ph) * np.cos(th) * v[0] - np.sin(ph) * np.sin(th) * v[1] + np.cos(ph) * v[2]
    z = np.cos(ph) * np.cos(th) * v[0] + np.cos(ph) * np.sin(th) * v[1] + np.sin(ph) * v[2]
    return x, y, z


def _enu2geo(ref, x, y, z):
    """SET.cpp:86-99 (vectorised, FP64 sqrt): ENU km -> geodetic."""
    lat_r, lon_r, d_r = ref
    xr = (_R_EARTH - d_r) * np.cos(lon_r / _RAD) * np.cos(lat_r / _RAD)
    yr = (_R_EARTH - d_r) * np.sin(lon_r / _RAD) * np.cos(lat_r / _RAD)
    zr = (_R_EARTH - d_r) * np.sin(lat_r / _RAD)
    ph, th = lat_r / _RAD, lon_r / _RAD
    xi = -np.sin(th) * x - np.sin(ph) * np.cos(th) * y + np.cos(ph) * np.cos(th) * z + xr
    yi = np.cos(th) * x - np.sin(ph) * np.sin(th) * y + np.cos(ph) * np.sin(th) * z + yr
    zi = np.cos(ph) * y + np.sin(ph) * z + zr
    r = np.sqrt(xi * xi + yi * yi + zi * zi)
    return _RAD * np.arcsin(zi / r), _RAD * np.arctan2(yi, xi), _R_EARTH - r


@dataclass
class TfmInputs:
    di: int
    dj: int
    dk: int
    fh: int
    fv: int
    bfr: int
    rc: np.ndarray        # RC[10], TFM.cpp:211
    sea: np.ndarray       # [di*dj] sea-level k
    topo: np.ndarray      # [di*dj] k units
    moho: np.ndarray      # [di*dj] rendered Moho k
    vel: np.ndarray       # [di*dj*dk] velocity, km/s
    vm: float
    ro: float
    dr: float

    @property
    def surf(self) -> np.ndarray:
        return self.topo + self.sea

    @property
    def n_fine(self) -> int:
        return (((self.di - 1) * self.fh + 1) * ((self.dj - 1) * self.fh + 1)
                * ((self.dk - 1) * self.fv + 1))


def _smooth2(seed, stream, di, dj, n_bumps, width):
    u = uniform(seed, stream, n_bumps * 4).reshape(n_bumps, 4)
    j, i = np.meshgrid(np.arange(dj), np.arange(di), indexing="ij")
    f = np.zeros((dj, di))
    for a, s_, cx, cy in u:
        w = width * (0.5 + s_)
        f += (2 * a - 1) * np.exp(-((i - cx * di) ** 2 + (j - cy * dj) ** 2) / (2 * w * w))
    return f.ravel()


def tfm_fm(seed: int = SEED0 + 6, di: int = 61, dj: int = 61, dk: int = 33, h: float = 10.0,
           fh: int = 1, fv: int = 2, bfr: int = 10, ref=(-15.25, -72.5)) -> TfmInputs:
    """Inputs of one gravity FM call at the reference's grid (TFM.cpp:190-225):
    sea[n] = k of sea level under column n (TFM.cpp:224), topography from 6
    seeded Gaussian ranges (<= 0.5 k units = 5 km, floored at -0.5, :225), Airy
    Moho M0 = sea - 35/h - 5 topo (:225) plus a +-0.5 k perturbation, mantle
    velocity 8 km/s +- 0.2 (3 seeded smooth layers), vm = 8, ro = -0.4
    (-drhomc), dr = 0.3 (drhov) (:195-196)."""
    rc = np.array([ref[0], ref[1], (dk - 3) * h / 2 - 5, (di - 1) * h / 2, (dj - 1) * h / 2,
                   (dk - 1) * h / 2, h, float(di), float(dj), float(dk)])
    j, i = np.meshgrid(np.arange(dj, dtype=np.float64), np.arange(di, dtype=np.float64),
                       indexing="ij")
    i, j = i.ravel(), j.ravel()
    lat, lon, _ = _enu2geo(rc[:3], i * h - rc[3], j * h - rc[4], (dk - 3.0) * h - rc[5])
    _, _, z = _geo2enu(rc[:3], lat, lon, np.zeros_like(lat))
    sea = (z + rc[5]) / h
    topo = np.maximum(0.5 * np.abs(_smooth2(seed, 1, di, dj, 6, 8.0)), -0.5)
    moho = sea - 35.0 / h - 5.0 * topo + 0.5 * np.tanh(_smooth2(seed, 2, di, dj, 8, 6.0))
    vel = np.empty(dk * dj * di)
    for k in range(dk):
        a = (k % 3) / 3.0
        vel[k * di * dj:(k + 1) * di * dj] = 8.0 + 0.2 * np.tanh(
            (1 - a) * _smooth2(seed, 10 + k // 3, di, dj, 8, 5.0)
            + a * _smooth2(seed, 11 + k // 3, di, dj, 8, 5.0))
    return TfmInputs(di, dj, dk, fh, fv, bfr, rc, sea, topo, moho, vel, 8.0, -0.4, 0.3)


WORKLOADS = {
    "pm_cpu_small": pm_cpu_small,
    "pm_blobs_1m": pm_blobs_1m,
    "adjoint_small": adjoint_small,
    "ensemble_256": ensemble_256,
    "uneven_48": uneven_48,
    "tfm_fm": tfm_fm,
}


# A Moho basis for the gravity-fit iteration on tfm_fm: six Gaussian bumps
# of the grid (k units, sigma 6 columns) centred in the interior; the step
# from the start (vm, ro, dr, 0..0) to the synthetic truth; the parameters'
# a-priori sigmas (LMA::fit's A = diag(1/sigma^2), TFM.cpp:160).
TFM_MOHO_CENTRES = [(20, 20), (30, 20), (40, 20), (20, 40), (30, 40), (40, 40)]
TFM_LM_STEP = np.array([0.05, 0.03, -0.03, 0.15, -0.10, 0.12, -0.08, 0.10, -0.12])
TFM_LM_SIGMA = np.array([0.1, 0.1, 0.1] + [0.3] * 6)


def tfm_moho_basis(w, centres=TFM_MOHO_CENTRES, sigma=6.0) -> np.ndarray:
    """[len(centres), di * dj] Gaussian bumps of the Moho grid."""
    j, i = np.meshgrid(np.arange(w.dj, dtype=np.float64), np.arange(w.di, dtype=np.float64),
                       indexing="ij")
    return np.stack([np.exp(-((i - ci) ** 2 + (j - cj) ** 2) / (2 * sigma ** 2)).ravel()
                     for ci, cj in centres])


# the reference's damping grid dmp = {{0,.3,1,3,10,100},{1e4,1e5}} as the
# (d0, d1) pairs LMA::fit iterates over (TFM.cpp:202,209)
GN_DAMPING = np.array([[a, b] for a in (0.0, 0.3, 1.0, 3.0, 10.0, 100.0) for b in (1e4, 1e5)])


@dataclass
class GnStep:
    J: np.ndarray          # [rows, n] float64, row-major (data x parameters)
    r: np.ndarray          # [rows] residuals
    sig: np.ndarray        # [n] a-priori parameter sigma
    b: np.ndarray          # [n] current parameters
    dm: np.ndarray         # [nd, 2] damping conditions (d0, d1)


def gn_step(seed: int = SEED0 + 7, rows: int = 32768, n: int = 1024) -> GnStep:
    """Inputs of one damped Gauss-Newton step (LMA::fit, TFM.cpp:157-161): a
    Jacobian whose columns span e^[-3,3] in scale, as event-location,
    velocity and Moho parameters do, with 30 % of its entries zero."""
    scale = np.exp(uniform(seed, 1, n, -3.0, 3.0))
    J = normal(seed, 2, rows * n).reshape(rows, n) * scale
    J[uniform(seed, 3, rows * n).reshape(rows, n) < 0.3] = 0.0
    return GnStep(J, normal(seed, 4, rows), uniform(seed, 5, n, 0.05, 4.0),
                  normal(seed, 6, n), GN_DAMPING.copy())
