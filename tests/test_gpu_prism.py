"""GPU parity of the prism cell formula against the FP64 oracle (oracle_prism_*).

FP64 path: exact closed form on every pair, <= 1e-11 (max-normalised).
FP32 path: point-mass far field (+ quadrupole for non-cubic cells), 3x3x3
Gauss-Legendre mid zone, FP64 closed-form near field: <= 1e-5.
FP64 zoned path (the FP32 path's algorithm in FP64): <= 1e-5 against the
oracle, <= 5e-6 against the FP32 path, transposes to 1e-12.
"""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def _t(a, dev, dt):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)


def stations_mixed(g, m, seed):
    """Stations above, beside, inside and below the grid, plus some exactly on
    cell centres, corners and faces."""
    c = g.centres()
    lo, hi = c.min(0), c.max(0)
    pad = np.array([abs(g.dx), abs(g.dy), abs(g.dz)])
    st = np.stack([W.uniform(seed, a, m, lo[a] - 3 * pad[a], hi[a] + 3 * pad[a])
                   for a in range(3)], axis=1)
    st[: m // 3, 2] = hi[2] + W.uniform(seed, 5, m // 3, 0.02, 2.0) * pad[2]   # above the top
    st[0] = c[len(c) // 2]                                                     # a cell centre
    st[1] = c[7] + 0.5 * np.array([g.dx, g.dy, g.dz])                          # a cell corner
    st[2] = c[11] + np.array([0.5 * g.dx, 0.0, 0.0])                           # a cell face
    return st


CASES = {
    "cubic": (W.Grid(24, 20, 12, 50.0, 50.0, -50.0, 100.0, 100.0, -100.0), 300),
    "aniso": (W.Grid(19, 23, 9, -300.0, 40.0, 10.0, 120.0, -80.0, -50.0), 257),
    "thin": (W.Grid(30, 6, 3, 0.0, 0.0, 0.0, 40.0, 40.0, 160.0), 64),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_prism_forward_vs_oracle(dev, orc, name):
    import torch
    g, m = CASES[name]
    print(f"[prism fwd] {name}", flush=True)
    st = stations_mixed(g, m, 3)
    rho = W.uniform(4, 1, g.n_cells, -400, 400)
    want = orc.prism_forward(g, W.KAPPA_SI, rho, st)
    want32 = orc.prism_forward(g, W.KAPPA_SI, rho.astype(np.float32), st)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI, formula=gfm.PRISM) as p:
        g64 = p.forward(_t(rho, dev, torch.float64)).cpu().numpy()
        g32 = p.forward(_t(rho, dev, torch.float32)).cpu().numpy()
    assert np.all(np.isfinite(g32)) and np.all(np.isfinite(g64))
    assert relerr(g64, want) < TOL64
    assert relerr(g32, want32) < TOL32


@pytest.mark.parametrize("name", sorted(CASES))
def test_prism_adjoint_vs_oracle(dev, orc, name):
    import torch
    g, m = CASES[name]
    print(f"[prism adj] {name}", flush=True)
    st = stations_mixed(g, m, 5)
    r = W.normal(6, 1, m, 1.0)
    want = orc.prism_adjoint(g, W.KAPPA_SI, st, r)
    want32 = orc.prism_adjoint(g, W.KAPPA_SI, st, r.astype(np.float32))
    with gfm.Plan(g, st, kappa=W.KAPPA_SI, formula=gfm.PRISM) as p:
        a64 = p.adjoint(_t(r, dev, torch.float64)).cpu().numpy()
        a32 = p.adjoint(_t(r, dev, torch.float32)).cpu().numpy()
    assert relerr(a64, want) < TOL64
    assert relerr(a32, want32) < TOL32


def test_prism_fp32_dot_product(dev):
    """The FP32 forward and adjoint correct the same pairs, so they stay
    transposes of each other."""
    import torch
    g, m = CASES["aniso"]
    st = stations_mixed(g, m, 7)
    rho = W.uniform(8, 1, g.n_cells, -1, 1)
    r = W.normal(8, 2, m, 1.0)
    with gfm.Plan(g, st, kappa=1.0, formula=gfm.PRISM) as p:
        gr = p.forward(_t(rho, dev, torch.float32)).cpu().numpy()
        a = p.adjoint(_t(r, dev, torch.float32)).cpu().numpy().astype(np.float64)
    rho32 = rho.astype(np.float32).astype(np.float64)
    r32 = r.astype(np.float32).astype(np.float64)
    lhs, rhs = gr @ r32, rho32 @ a
    assert abs(lhs - rhs) <= 1e-6 * (np.abs(gr) @ np.abs(r32))


def test_prism_blobs_fp32_vs_fp64_gpu(dev):
    """Half-size pm_blobs grid (64x64x32 cells of 1 km, 4,096 stations with
    topography): the FP32 hybrid against the FP64 closed form everywhere."""
    import torch
    w = W.pm_blobs_1m(n=64, nz=32, h=1000.0)
    r = W.normal(9, 1, w.n_stations, 0.5)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        g64 = p.forward(_t(w.rho, dev, torch.float64)).cpu().numpy()
        g32 = p.forward(_t(w.rho, dev, torch.float32)).cpu().numpy()
        a64 = p.adjoint(_t(r, dev, torch.float64)).cpu().numpy()
        a32 = p.adjoint(_t(r, dev, torch.float32)).cpu().numpy()
    assert relerr(g32, g64) < TOL32
    assert relerr(a32, a64) < TOL32


# ZTM_FP64_ZONED: the FP32 path's algorithm (zones, quadrupole far field,
# Gauss-Legendre mid zone, closed-form near box) in FP64. Against the exact
# oracle it carries the FP32 path's truncation error (TOL32); against the
# FP32 path only FP32 rounding separates them (2.4e-6 of the maximum measured
# on the thin grid's adjoint, whose 40 x 40 x 160 m cells put stations a
# fraction of a cell from terms far larger than the sums they cancel into);
# and its forward and adjoint correct the same pairs, so they are transposes
# to FP64 rounding.
TOL_SAME_ALG = 5e-6


@pytest.mark.parametrize("name", sorted(CASES))
def test_prism_zoned_vs_oracle_and_fp32(dev, orc, name):
    import torch
    g, m = CASES[name]
    print(f"[prism zoned] {name}", flush=True)
    st = stations_mixed(g, m, 11)
    rho = W.uniform(12, 1, g.n_cells, -400, 400)
    r = W.normal(12, 2, m, 1.0)
    want_f = orc.prism_forward(g, W.KAPPA_SI, rho, st)
    want_a = orc.prism_adjoint(g, W.KAPPA_SI, st, r)
    rho32 = rho.astype(np.float32)
    r32 = r.astype(np.float32)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI, formula=gfm.PRISM) as p:
        gz = p.forward(_t(rho, dev, torch.float64), zoned=True).cpu().numpy()
        az = p.adjoint(_t(r, dev, torch.float64), zoned=True).cpu().numpy()
        gz32 = p.forward(_t(rho32, dev, torch.float64), zoned=True).cpu().numpy()
        az32 = p.adjoint(_t(r32, dev, torch.float64), zoned=True).cpu().numpy()
        g32 = p.forward(_t(rho32, dev, torch.float32)).cpu().numpy()
        a32 = p.adjoint(_t(r32, dev, torch.float32)).cpu().numpy()
        g64 = p.forward(_t(rho, dev, torch.float64)).cpu().numpy()
    assert np.all(np.isfinite(gz)) and np.all(np.isfinite(az))
    assert relerr(gz, want_f) < TOL32
    assert relerr(az, want_a) < TOL32
    assert relerr(g32, gz32) < TOL_SAME_ALG
    assert relerr(a32, az32) < TOL_SAME_ALG
    # the zones change the result: the zoned path is not the exact one
    assert relerr(gz, g64) > 1e-10


@pytest.mark.parametrize("name", sorted(CASES))
def test_prism_zoned_dot_product(dev, name):
    import torch
    g, m = CASES[name]
    st = stations_mixed(g, m, 13)
    rho = W.uniform(14, 1, g.n_cells, -1, 1)
    r = W.normal(14, 2, m, 1.0)
    with gfm.Plan(g, st, kappa=1.0, formula=gfm.PRISM) as p:
        gr = p.forward(_t(rho, dev, torch.float64), zoned=True).cpu().numpy()
        a = p.adjoint(_t(r, dev, torch.float64), zoned=True).cpu().numpy()
    lhs, rhs = gr @ r, rho @ a
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(gr) @ np.abs(r))


def test_prism_zoned_blobs(dev):
    """Half-size pm_blobs grid: zoned FP64 against the exact FP64 closed form
    (truncation of the zones) and against the FP32 hybrid (rounding only)."""
    import torch
    w = W.pm_blobs_1m(n=64, nz=32, h=1000.0)
    r = W.normal(15, 1, w.n_stations, 0.5)
    rho32 = w.rho.astype(np.float32)
    r32 = r.astype(np.float32)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        g64 = p.forward(_t(w.rho, dev, torch.float64)).cpu().numpy()
        gz = p.forward(_t(w.rho, dev, torch.float64), zoned=True).cpu().numpy()
        gz32 = p.forward(_t(rho32, dev, torch.float64), zoned=True).cpu().numpy()
        g32 = p.forward(_t(rho32, dev, torch.float32)).cpu().numpy()
        a64 = p.adjoint(_t(r, dev, torch.float64)).cpu().numpy()
        az = p.adjoint(_t(r, dev, torch.float64), zoned=True).cpu().numpy()
        az32 = p.adjoint(_t(r32, dev, torch.float64), zoned=True).cpu().numpy()
        a32 = p.adjoint(_t(r32, dev, torch.float32)).cpu().numpy()
    assert relerr(gz, g64) < TOL32
    assert relerr(az, a64) < TOL32
    assert relerr(g32, gz32) < TOL_SAME_ALG
    assert relerr(a32, az32) < TOL_SAME_ALG


def test_zoned_point_mass_is_fp64(dev):
    """Point-mass plans have no zones: ZTM_FP64_ZONED is the FP64 path."""
    import torch
    g, m = CASES["aniso"]
    st = stations_mixed(g, m, 16)
    rho = W.uniform(17, 1, g.n_cells, -1, 1)
    with gfm.Plan(g, st, kappa=1.0) as p:
        a = p.forward(_t(rho, dev, torch.float64)).cpu().numpy()
        b = p.forward(_t(rho, dev, torch.float64), zoned=True).cpu().numpy()
    assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        p2 = gfm.Plan(g, st, kappa=1.0)
        try:
            p2.forward(_t(rho, dev, torch.float32), zoned=True)
        finally:
            p2.close()
