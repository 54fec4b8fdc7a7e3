"""GPU parity of the point-mass forward/adjoint kernels against the FP64 oracle.

Tolerances (max |gpu - oracle| / max |oracle|):
  FP32 path  <= 1e-6  (the product target is 1e-5; typical is ~1e-7)
  FP64 path  <= 1e-12
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu

TOL32 = 1e-6
TOL64 = 1e-12


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def ref_case(seed=7, di=61, dj=61, dk=33, fh=1, fv=2):
    return W.tfm_shaped(seed, di, dj, dk, fh, fv)


# ------------------------------------------------------------ reference ABI
@pytest.mark.parametrize("geom", [(61, 61, 33, 1, 2, 10), (23, 17, 6, 1, 2, 3),
                                  (12, 9, 4, 2, 3, 2), (8, 8, 3, 1, 1, 4),
                                  (9, 6, 2, 3, 1, 1)])
def test_gf_matches_oracle(dev, orc, geom):
    import torch
    di, dj, dk, fh, fv, bfr = geom
    print(f"[gf] geom di={di} dj={dj} dk={dk} fh={fh} fv={fv} bfr={bfr}", flush=True)
    rho, s = ref_case(di=di, dj=dj, dk=dk, fh=fh, fv=fv)
    g0 = W.uniform(11, 5, di * dj, -1, 1).astype(np.float32)     # g is accumulated into
    want = orc.gf(bfr, di, dj, dk, fh, fv, rho, s, g=g0.astype(np.float64))
    g = torch.from_numpy(g0.copy()).to(dev)
    gfm.GF(bfr, di, dj, dk, fh, fv, g, torch.from_numpy(rho).to(dev),
           torch.from_numpy(s).to(dev))
    got = g.cpu().numpy()
    interior = np.zeros((dj, di), bool)
    interior[bfr:dj - bfr, bfr:di - bfr] = True
    interior = interior.ravel()
    assert np.array_equal(got[~interior], g0[~interior]), "exterior stations must be untouched"
    if interior.any():
        assert relerr(got[interior] - g0[interior], want[interior] - g0[interior]) < TOL32


def test_nan_cells_are_skipped(dev, orc):
    """NaN densities: GF's test fabs(rho) > .001f is false for NaN
    (GFM_GPU.cu:19), so the reference skips those cells. ztm_gf, the plan
    forward (FP32 and FP64, threshold 0 and .001) and the batched forward
    (MFMA split and the up-to-4-model pass) must give exactly the result with
    those cells set to 0, and that result must match the oracle."""
    import torch
    di, dj, dk, fh, fv, bfr = 23, 17, 6, 1, 2, 3
    rho, s = ref_case(seed=13, di=di, dj=dj, dk=dk, fh=fh, fv=fv)
    nan_at = np.unique((W.uniform(13, 9, 40) * rho.size).astype(int))
    rn = rho.copy()
    rn[nan_at] = np.nan
    r0 = rho.copy()
    r0[nan_at] = 0.0
    print(f"[nan] gf {di}x{dj}x{dk} nan cells {nan_at.size}", flush=True)
    want = orc.gf(bfr, di, dj, dk, fh, fv, rn, s)
    assert np.array_equal(want, orc.gf(bfr, di, dj, dk, fh, fv, r0, s))
    outs = []
    for field in (rn, r0):
        g = torch.zeros(di * dj, dtype=torch.float32, device=dev)
        gfm.GF(bfr, di, dj, dk, fh, fv, g, torch.from_numpy(field).to(dev),
               torch.from_numpy(s).to(dev))
        outs.append(g.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    assert relerr(outs[0], want) < TOL32
    w = W.pm_cpu_small()
    rn = w.rho.copy()
    cells = np.unique((W.uniform(13, 10, 64) * rn.size).astype(int))
    rn[cells] = np.nan
    r0 = w.rho.copy()
    r0[cells] = 0.0
    for thr in (0.0, 1e-3):
        print(f"[nan] plan forward threshold {thr}", flush=True)
        with gfm.Plan(w.grid, w.stations, kappa=w.kappa, threshold=thr) as p:
            for dt in (torch.float32, torch.float64):
                a = p.forward(_torch(rn, dev, dt)).cpu().numpy()
                b = p.forward(_torch(r0, dev, dt)).cpu().numpy()
                assert np.array_equal(a, b), (thr, dt)
            assert relerr(a, orc.pm_forward(w.grid, w.kappa, r0, w.stations)) < TOL64
            for B in (3, 40):
                P = np.stack([rn] * B)
                Y = p.forward_batched(_torch(P, dev, torch.float32)).cpu().numpy()
                Y0 = p.forward_batched(_torch(np.stack([r0] * B), dev, torch.float32))
                assert np.array_equal(Y, Y0.cpu().numpy()), (thr, B)


def test_gfm_host_and_program(dev, orc):
    di, dj, dk, fh, fv, bfr = 61, 61, 33, 1, 2, 10
    rho, s = ref_case(seed=9)
    want = orc.gf(bfr, di, dj, dk, fh, fv, rho, s)
    got = gfm.gfm_host(bfr, di, dj, dk, fh, fv, rho, s)
    assert relerr(got, want) < TOL32
    exe = os.path.join(os.path.dirname(gfm.LIB_PATH), "bin", "GFM")
    with tempfile.TemporaryDirectory() as d:
        rho.tofile(os.path.join(d, "R.dat"))
        s.tofile(os.path.join(d, "S.dat"))
        env = dict(os.environ, ZTM_GFM_DIR=d)
        p = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        g = np.fromfile(os.path.join(d, "G.dat"), dtype=np.float32)
    assert g.size == di * dj
    assert np.array_equal(g, got), "GFM program and ztm_gfm_host must agree bitwise"


# ------------------------------------------------------------ general plans
def _torch(a, dev, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


@pytest.mark.parametrize("st_per_lane", [1, 2, 4, 8])
def test_forward_small_vs_oracle(dev, orc, st_per_lane):
    import torch
    w = W.pm_cpu_small()
    print(f"[fwd] {w.name} S={st_per_lane}", flush=True)
    want = orc.pm_forward(w.grid, w.kappa, w.rho, w.stations)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_tuning(st_per_lane=st_per_lane)
        g32 = p.forward(_torch(w.rho, dev, torch.float32)).cpu().numpy()
        g64 = p.forward(_torch(w.rho, dev, torch.float64)).cpu().numpy()
    # FP32 path vs oracle on the FP32-rounded density (kernel error only)
    want32 = orc.pm_forward(w.grid, w.kappa, w.rho.astype(np.float32), w.stations)
    assert relerr(g32, want32) < TOL32
    assert relerr(g32, want) < 1e-5
    assert relerr(g64, want) < TOL64


@pytest.mark.parametrize("cells_per_lane", [4, 8, 16])
def test_adjoint_small_vs_oracle(dev, orc, cells_per_lane):
    import torch
    w = W.adjoint_small()
    r = w.extra["residual"]
    print(f"[adj] {w.name} C={cells_per_lane}", flush=True)
    want = orc.pm_adjoint(w.grid, w.kappa, w.stations, r)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_tuning(cells_per_lane=cells_per_lane)
        a32 = p.adjoint(_torch(r, dev, torch.float32)).cpu().numpy()
        a64 = p.adjoint(_torch(r, dev, torch.float64)).cpu().numpy()
    want32 = orc.pm_adjoint(w.grid, w.kappa, w.stations, r.astype(np.float32))
    assert relerr(a32, want32) < TOL32
    assert relerr(a64, want) < TOL64


@pytest.mark.parametrize("cells_per_lane", [8, 16])
@pytest.mark.parametrize("shape", [(128, 18, 3, 500), (200, 20, 2, 700), (121, 16, 2, 300)])
def test_adjoint_row_table_shapes(dev, orc, shape, cells_per_lane):
    """Shapes whose adjoint workgroups mix the row-term table path (segments
    in one k plane over at most 16 rows) and the general path, with ragged
    row groups and lanes past the last segment; stations inside the grid
    exercise the clamped rows and the near field."""
    import torch
    nx, ny, nz, m = shape
    print(f"[adj-table] {shape} C={cells_per_lane}", flush=True)
    g = W.Grid(nx, ny, nz, 100.0, -50.0, 30.0, 110.0, -80.0, -60.0)
    c = g.centres()
    lo, hi = c.min(0) - 150, c.max(0) + 150
    st = np.stack([W.uniform(11, a, m, lo[a], hi[a]) for a in range(3)], axis=1)
    r = W.normal(11, 4, m, 1.0)
    want = orc.pm_adjoint(g, W.KAPPA_SI, st, r)
    want32 = orc.pm_adjoint(g, W.KAPPA_SI, st, r.astype(np.float32))
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        p.set_tuning(cells_per_lane=cells_per_lane)
        a32 = p.adjoint(_torch(r, dev, torch.float32)).cpu().numpy()
        a64 = p.adjoint(_torch(r, dev, torch.float64)).cpu().numpy()
    assert np.all(np.isfinite(a32))
    assert relerr(a32, want32) < TOL32
    assert relerr(a64, want) < TOL64


@pytest.mark.parametrize("shape", [(37, 5, 3, 13), (1, 1, 1, 1), (5, 1, 7, 600),
                                   (130, 3, 2, 1030), (64, 64, 1, 70)])
def test_ragged_and_anisotropic(dev, orc, shape):
    """Odd sizes, non-cubic cells, negative spacing, stations inside the grid."""
    import torch
    nx, ny, nz, m = shape
    print(f"[ragged] {shape}", flush=True)
    g = W.Grid(nx, ny, nz, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0)
    c = g.centres()
    lo, hi = c.min(0) - 100, c.max(0) + 100
    st = np.stack([W.uniform(3, a, m, lo[a], hi[a]) for a in range(3)], axis=1)
    st[0] = c[min(3, len(c) - 1)]                   # coincident with a cell centre
    rho = W.uniform(5, 1, g.n_cells, -500, 500)
    r = W.normal(5, 2, m, 1.0)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI, threshold=50.0) as p:
        for prec, dt, tol in ((gfm.FP32, torch.float32, TOL32), (gfm.FP64, torch.float64, TOL64)):
            rr = rho.astype(np.float32) if prec == gfm.FP32 else rho
            want = orc.pm_forward(g, W.KAPPA_SI, rr, st, threshold=50.0)
            got = p.forward(_torch(rho, dev, dt)).cpu().numpy()
            assert np.all(np.isfinite(got))
            assert relerr(got, want) < tol
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as q:
            for prec, dt, tol in ((gfm.FP32, torch.float32, TOL32),
                                  (gfm.FP64, torch.float64, TOL64)):
                rr = r.astype(np.float32) if prec == gfm.FP32 else r
                want = orc.pm_adjoint(g, W.KAPPA_SI, st, rr)
                got = q.adjoint(_torch(r, dev, dt)).cpu().numpy()
                assert np.all(np.isfinite(got))
                assert relerr(got, want) < tol


@pytest.mark.parametrize("formula", [gfm.POINT_MASS, gfm.PRISM])
def test_near_singular_stations(dev, orc, formula):
    """Stations 1e-5..0.49 cells from cell centres: the FP32 path must not let the
    near-singular pair swamp the rest of its row (per-station error <= 1e-5 of
    max(|g_m|, median |g|))."""
    import torch
    g = W.Grid(24, 20, 10, 0.0, 0.0, 0.0, 1.0, 1.0, -0.5)
    c = g.centres()
    offs = [1e-5, 1e-4, 1e-3, 1e-2, 0.1, 0.3, 0.49, 0.0]
    dirs = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [0.3, 0.4, 0.5], [0, 1, 0.2]], float)
    st = []
    for t, o in enumerate(offs):
        for d in dirs:
            cell = c[(37 * t + 11 * len(st)) % len(c)]
            st.append(cell + o * d / np.linalg.norm(d))
    st = np.array(st)
    rho = W.uniform(21, 1, g.n_cells, -300, 300)
    r = W.normal(21, 2, len(st), 1.0)
    fwd = orc.pm_forward if formula == gfm.POINT_MASS else orc.prism_forward
    adj = orc.pm_adjoint if formula == gfm.POINT_MASS else orc.prism_adjoint
    want = fwd(g, W.KAPPA_SI, rho.astype(np.float32), st)
    want_a = adj(g, W.KAPPA_SI, st, r.astype(np.float32))
    with gfm.Plan(g, st, kappa=W.KAPPA_SI, formula=formula) as p:
        got = p.forward(_torch(rho, dev, torch.float32)).cpu().numpy()
        got_a = p.adjoint(_torch(r, dev, torch.float32)).cpu().numpy()
    # per element: 1e-5 for the point mass; the prism's far-field multipole
    # term leaves up to ~2e-5 of a single cell's contribution just beyond its
    # mid zone, so 5e-5 there; max-normalised (the product metric) 1e-5 for both
    tol = 1e-5 if formula == gfm.POINT_MASS else 5e-5
    scale = np.maximum(np.abs(want), np.median(np.abs(want)))
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - want) / scale) < tol
    assert relerr(got, want) < 1e-5
    scale_a = np.maximum(np.abs(want_a), np.median(np.abs(want_a)))
    assert np.max(np.abs(got_a - want_a) / scale_a) < tol
    assert relerr(got_a, want_a) < 1e-5


def test_zero_density_and_no_stations(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        g = p.forward(torch.zeros(w.grid.n_cells, device=dev))
        assert torch.count_nonzero(g).item() == 0
    with gfm.Plan(w.grid, np.zeros((0, 3)), kappa=w.kappa) as p:
        out = torch.full((w.grid.n_cells,), 7.0, device=dev)
        p.adjoint(torch.zeros(0, device=dev), out=out)
        assert torch.count_nonzero(out).item() == 0


def test_deterministic(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        rho = _torch(w.rho, dev, torch.float32)
        a, b = p.forward(rho), p.forward(rho)
        assert torch.equal(a, b)
        r = torch.randn(w.n_stations, device=dev)
        assert torch.equal(p.adjoint(r), p.adjoint(r))


def test_errors_are_loud(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        with pytest.raises(ValueError):
            p.forward(torch.zeros(5, device=dev))
        with pytest.raises(gfm.ZtmError):
            p.forward_ptr(7, 0, 0, None)
        with pytest.raises(gfm.ZtmError):
            p.set_tuning(st_per_lane=3)
    with pytest.raises(gfm.ZtmError):
        gfm.gfm_host(10, 61, 61, 33, 1, 2, np.zeros(100, np.float32), np.zeros(61 * 61))


# ------------------------------------------------------------ full size
@pytest.fixture(scope="module")
def blobs():
    return W.pm_blobs_1m()


def test_blobs_1m_forward(dev, orc, blobs):
    """1,048,576 cells x 16,384 stations: FP32 vs FP64 GPU everywhere, both vs the
    oracle on a seeded station sample."""
    import torch
    w = blobs
    print("[full] pm_blobs_1m forward", flush=True)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        g32 = p.forward(_torch(w.rho, dev, torch.float32)).cpu().numpy()
        g64 = p.forward(_torch(w.rho, dev, torch.float64)).cpu().numpy()
    assert relerr(g32, g64) < 1e-5
    idx = np.unique((W.uniform(77, 1, 48) * w.n_stations).astype(int))
    want = orc.pm_forward(w.grid, w.kappa, w.rho, w.stations[idx])
    assert relerr(g64[idx], want) < 1e-11
    assert relerr(g32[idx], want) < 1e-5


def test_blobs_1m_adjoint_dot_product(dev, blobs):
    """<G rho, r> = <rho, G^T r> at full size (FP64 GPU), FP32 adjoint vs FP64."""
    import torch
    w = blobs
    print("[full] pm_blobs_1m adjoint", flush=True)
    r = W.normal(5, 9, w.n_stations, 0.5)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        g64 = p.forward(_torch(w.rho, dev, torch.float64)).cpu().numpy()
        a64 = p.adjoint(_torch(r, dev, torch.float64)).cpu().numpy()
        a32 = p.adjoint(_torch(r, dev, torch.float32)).cpu().numpy()
    lhs, rhs = float(g64 @ r), float(w.rho @ a64)
    assert abs(lhs - rhs) <= 1e-10 * (np.abs(g64) @ np.abs(r))
    assert relerr(a32, a64) < 1e-5
