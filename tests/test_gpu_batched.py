"""Batched (ensemble) forward on split-BF16 MFMA against the single-model FP32/FP64
kernels and the FP64 oracle. Tolerance: 1e-5 of max |g| per model (the product
target); typical is ~1e-6."""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def models(seed, B, n, lo=-300.0, hi=300.0):
    return np.stack([W.uniform(seed, 200 + b, n, lo, hi) for b in range(B)]).astype(np.float32)


def test_batched_small_vs_single_and_oracle(dev, orc):
    import torch
    w = W.pm_cpu_small()
    P = models(3, 37, w.grid.n_cells)                      # 37 -> padded to 2 tiles of 32
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
        singles = [p.forward(torch.from_numpy(P[b]).to(dev)).cpu().numpy() for b in range(37)]
    for b in range(37):
        assert relerr(Y[b], singles[b]) < 1e-5, b
    for b in (0, 36):
        want = orc.pm_forward(w.grid, w.kappa, P[b].astype(np.float64), w.stations)
        assert relerr(Y[b], want) < 1e-5


def test_batched_ragged(dev, orc):
    """N = 555 (not a multiple of 16, rows of 37 cells so chunks cross rows),
    13 stations (one partial tile), 5 models, negative spacings."""
    import torch
    g = W.Grid(37, 5, 3, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0)
    c = g.centres()
    st = np.stack([W.uniform(4, a, 13, c[:, a].min() - 200, c[:, a].max() + 200)
                   for a in range(2)] + [W.uniform(4, 2, 13, 80.0, 300.0)], axis=1)
    P = models(5, 5, g.n_cells)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
    for b in range(5):
        want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st)
        assert relerr(Y[b], want) < 2e-6, b


def test_batched_256_half_size(dev):
    """256 models, 48x48x24 cells of 1 km, 2,304 stations at z = +10 m: both
    splits and a long-slice setting against the FP64 single-model kernel on 4
    models; repeated runs are bitwise equal."""
    import torch
    w = W.ensemble_256(n=48, nz=24)
    Pt = torch.from_numpy(w.rho).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Y = p.forward_batched(Pt).cpu().numpy()            # default: 3-way split
        Y2 = p.forward_batched(Pt).cpu().numpy()
        p.set_batched_tuning(split=2)
        Yh = p.forward_batched(Pt).cpu().numpy()
        p.set_batched_tuning(slices=3, split=3)
        Y3 = p.forward_batched(Pt).cpu().numpy()
        for b in (0, 85, 170, 255):
            g64 = p.forward(torch.from_numpy(w.rho[b].astype(np.float64)).to(dev)).cpu().numpy()
            assert relerr(Y[b], g64) < 2e-6, b
            assert relerr(Yh[b], g64) < 1e-5, b
            assert relerr(Y3[b], g64) < 1e-5, b
    assert np.array_equal(Y, Y2)


def test_batched_rejects(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        with pytest.raises(gfm.ZtmError):
            p.forward_batched(torch.zeros(2, w.grid.n_cells, device=dev))
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        with pytest.raises(gfm.ZtmError):
            p.forward_batched(torch.zeros(257, w.grid.n_cells, device=dev))


def test_batched_narrow_rows(dev, orc):
    """nx < 8 (a lane's 8 cells cross several rows: the generic stepping
    path), nx = 8 (the two-row path at its limit), against the oracle."""
    import torch
    for nx, ny, nz in ((5, 7, 9), (3, 4, 20), (8, 6, 5)):
        g = W.Grid(nx, ny, nz, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0)
        c = g.centres()
        st = np.stack([W.uniform(6, a, 40, c[:, a].min() - 50, c[:, a].max() + 50)
                       for a in range(2)] + [W.uniform(6, 2, 40, 20.0, 200.0)], axis=1)
        P = models(7 + nx, 9, g.n_cells)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
        for b in (0, 4, 8):
            want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st)
            assert relerr(Y[b], want) < 2e-6, (nx, b)
