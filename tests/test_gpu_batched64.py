"""FP64 ensemble (G generated once per tile, reused across the batch on FP64
MFMA) against the FP64 oracle and the FP64 single-model kernels. G is the
same exact point mass on every pair in all three, so only the order of the
FP64 sums differs: 1e-12 of max |.| per model."""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def models(seed, B, n, lo=-300.0, hi=300.0):
    return np.stack([W.uniform(seed, 200 + b, n, lo, hi) for b in range(B)])


def residuals(seed, B, m, sigma=0.5):
    return np.stack([W.normal(seed, 300 + b, m, sigma) for b in range(B)])


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


@pytest.mark.parametrize("B", [1, 5, 16, 37, 256])
def test_batched64_small_vs_oracle_and_single(dev, orc, B):
    """pm_cpu_small: every model-tile count (1 .. 16 tiles of 16 columns), a
    partial last tile, against the oracle on the first and last model and the
    FP64 single-model kernels on all of them."""
    w = W.pm_cpu_small()
    print(f"[batched64 small] B={B}", flush=True)
    P = models(3, B, w.grid.n_cells)
    R = residuals(4, B, w.n_stations)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Y = p.forward_batched64(_t(P, dev)).cpu().numpy()
        X = p.adjoint_batched64(_t(R, dev)).cpu().numpy()
        for b in range(B):
            assert relerr(Y[b], p.forward(_t(P[b], dev)).cpu().numpy()) < TOL, b
            assert relerr(X[b], p.adjoint(_t(R[b], dev)).cpu().numpy()) < TOL, b
    for b in sorted({0, B - 1}):
        assert relerr(Y[b], orc.pm_forward(w.grid, w.kappa, P[b], w.stations)) < TOL, b
        assert relerr(X[b], orc.pm_adjoint(w.grid, w.kappa, w.stations, R[b])) < TOL, b


def test_batched64_ragged(dev, orc):
    """N = 555 (rows of 37 cells, chunks cross rows), 13 and 40 stations,
    nx < 8 grids, negative spacings."""
    cases = [((37, 5, 3, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0), 13),
             ((37, 5, 3, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0), 40),
             ((5, 7, 9, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0), 40),
             ((3, 4, 20, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0), 77)]
    for gi, (gargs, m) in enumerate(cases):
        g = W.Grid(*gargs)
        c = g.centres()
        st = np.stack([W.uniform(8 + gi, a, m, c[:, a].min() - 200, c[:, a].max() + 200)
                       for a in range(2)] + [W.uniform(8 + gi, 2, m, 80.0, 300.0)], axis=1)
        P = models(5 + gi, 7, g.n_cells)
        R = residuals(12 + gi, 7, m)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            Y = p.forward_batched64(_t(P, dev)).cpu().numpy()
            X = p.adjoint_batched64(_t(R, dev)).cpu().numpy()
        for b in range(7):
            assert relerr(Y[b], orc.pm_forward(g, W.KAPPA_SI, P[b], st)) < TOL, (gi, b)
            assert relerr(X[b], orc.pm_adjoint(g, W.KAPPA_SI, st, R[b])) < TOL, (gi, b)


def test_batched64_threshold_and_nan(dev, orc):
    """The plan's threshold and NaN skipping apply to the models, as in the
    single-model forward: cells with |rho| <= thr or NaN contribute nothing."""
    w = W.pm_cpu_small()
    P = models(6, 9, w.grid.n_cells, -1.0, 1.0)
    P[:, ::7] = np.nan
    P[:, 3::11] = 0.25                       # at the threshold: skipped
    clean = np.where(np.isnan(P) | (np.abs(P) <= 0.25), 0.0, P)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, threshold=0.25) as p:
        Y = p.forward_batched64(_t(P, dev)).cpu().numpy()
    assert np.all(np.isfinite(Y))
    for b in (0, 8):
        want = orc.pm_forward(w.grid, w.kappa, clean[b], w.stations)
        assert relerr(Y[b], want) < TOL, b


def test_batched64_dot_product_and_direct_adjoint(dev):
    """64 x 64 x 16 cells (512 cell tiles: the adjoint writes X directly,
    one K-slice) and 48 x 48 x 8 (several K-slices and the reduce): the
    forward and adjoint are transposes, <G P, R> = <P, G^T R> to 1e-12."""
    for n, nz in ((64, 16), (48, 8)):
        w = W.ensemble_256(n=n, nz=nz)
        B = 40
        P = w.rho[:B].astype(np.float64)
        R = residuals(16, B, w.n_stations)
        with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
            Y = p.forward_batched64(_t(P, dev)).cpu().numpy()
            X = p.adjoint_batched64(_t(R, dev)).cpu().numpy()
            for b in (0, B - 1):
                assert relerr(X[b], p.adjoint(_t(R[b], dev)).cpu().numpy()) < TOL, (n, b)
        lhs = np.einsum("bm,bm->b", Y, R)
        rhs = np.einsum("bn,bn->b", P, X)
        scale = np.einsum("bm,bm->b", np.abs(Y), np.abs(R))
        assert np.max(np.abs(lhs - rhs) / scale) < TOL, n


def test_batched64_rejects_and_empty(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        with pytest.raises(gfm.ZtmError):
            p.forward_batched64(_t(np.zeros((2, w.grid.n_cells)), dev))
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        with pytest.raises(gfm.ZtmError):
            p.adjoint_batched64(_t(np.zeros((257, w.n_stations)), dev))
        with pytest.raises(ValueError):
            p.forward_batched64(torch.zeros(2, w.grid.n_cells, device=dev))   # float32
        assert p.forward_batched64(_t(np.zeros((0, w.grid.n_cells)), dev)).shape == \
            (0, w.n_stations)
    with gfm.Plan(w.grid, np.zeros((0, 3)), kappa=w.kappa) as p:
        out = torch.full((5, w.grid.n_cells), 7.0, dtype=torch.float64, device=dev)
        p.adjoint_batched64(torch.zeros(5, 0, dtype=torch.float64, device=dev), out=out)
        assert not torch.any(out).item()


def test_batched64_full_size(dev, orc):
    """ensemble_256 at its full size (96x96x48 cells, 9,216 stations, 256
    models and vectors): models / vectors 0, 85, 170, 255 against the oracle
    on a seeded sample (64 stations; 2,048 cells) and against the FP64
    single-model kernels on 0 and 255 everywhere."""
    w = W.ensemble_256()
    R = residuals(14, 256, w.n_stations)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Y = p.forward_batched64(_t(w.rho, dev)).cpu().numpy()
        X = p.adjoint_batched64(_t(R, dev)).cpu().numpy()
        for b in (0, 255):
            g64 = p.forward(_t(w.rho[b], dev)).cpu().numpy()
            a64 = p.adjoint(_t(R[b], dev)).cpu().numpy()
            assert relerr(Y[b], g64) < TOL, b
            assert relerr(X[b], a64) < TOL, b
    st = np.unique((W.uniform(15, 1, 64) * w.n_stations).astype(np.int64))
    cells = np.unique((W.uniform(15, 2, 2048) * w.grid.n_cells).astype(np.int64))
    for b in (0, 85, 170, 255):
        g_or = orc.pm_forward(w.grid, w.kappa, w.rho[b].astype(np.float64), w.stations[st])
        a_or = orc.pm_adjoint_cells(w.grid, w.kappa, w.stations, R[b], cells)
        assert relerr(Y[b][st], g_or) < TOL, b
        assert relerr(X[b][cells], a_or) < TOL, b
