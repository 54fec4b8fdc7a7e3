"""Batched (ensemble) forward and adjoint on split-BF16 and split-FP16 MFMA
against the single-model FP32/FP64 kernels and the FP64 oracle. Tolerance:
1e-5 of max |g| per model (the product target); typical is ~1e-6."""
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
    """256 models, 48x48x24 cells of 1 km, 2,304 stations at z = +10 m: the
    default (FP16 hi/lo), BF16 2-way and 3-way splits and a long-slice setting
    against the FP64 single-model kernel on 4 models; repeated runs are
    bitwise equal."""
    import torch
    w = W.ensemble_256(n=48, nz=24)
    Pt = torch.from_numpy(w.rho).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        assert p.batched_tuning() == (0, 4, 4)             # stations at one height
        Y = p.forward_batched(Pt).cpu().numpy()            # default: FP16 hi/lo split
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


# ------------------------------------------------------------ batched adjoint
# G^T R on the same split-BF16 scheme. The 3-way split is held to the product
# bar (1e-5 of max |G^T r| per vector); hi/lo (~2^-16 per entry) to 1e-4,
# since zero-mean residuals cancel in the sum while the entry errors do not.

def residuals(seed, B, m, sigma=0.5):
    return np.stack([W.normal(seed, 300 + b, m, sigma) for b in range(B)]).astype(np.float32)


def test_batched_adjoint_small_vs_single_and_oracle(dev, orc):
    import torch
    w = W.pm_cpu_small()
    R = residuals(11, 37, w.n_stations)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
        singles = [p.adjoint(torch.from_numpy(R[b].astype(np.float64)).to(dev)).cpu().numpy()
                   for b in range(37)]
    assert X.shape == (37, w.grid.n_cells) and X.dtype == np.float32
    for b in range(37):
        assert relerr(X[b], singles[b]) < 1e-5, b
    for b in (0, 36):
        want = orc.pm_adjoint(w.grid, w.kappa, w.stations, R[b].astype(np.float64))
        assert relerr(X[b], want) < 1e-5, b


def test_batched_adjoint_ragged(dev, orc):
    """N = 555 (a partial tile of 128 cells), 13 and 40 stations (partial
    station chunks, one or three chunks), 5 vectors, negative spacings; and
    nx < 8 grids."""
    import torch
    cases = [((37, 5, 3, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0), 13),
             ((37, 5, 3, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0), 40),
             ((5, 7, 9, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0), 40),
             ((3, 4, 20, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0), 77)]
    for gi, (gargs, m) in enumerate(cases):
        g = W.Grid(*gargs)
        c = g.centres()
        st = np.stack([W.uniform(8 + gi, a, m, c[:, a].min() - 200, c[:, a].max() + 200)
                       for a in range(2)] + [W.uniform(8 + gi, 2, m, 80.0, 300.0)], axis=1)
        R = residuals(12 + gi, 5, m)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
        for b in range(5):
            want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
            assert relerr(X[b], want) < 1e-5, (gi, b)


@pytest.mark.parametrize("n,nz", [(48, 24), (64, 12)])
def test_batched_adjoint_256_half_size(dev, n, nz):
    """256 vectors, n x n x nz cells of 1 km, n^2 stations (144 / 256 station
    chunks in one K-slice, stored straight from the accumulators with 16-byte
    stores; nx = 64 takes the shared-row generation): both splits against the
    FP64 single-vector
    adjoint on 4 vectors, bitwise repeatability, and the dot-product identity
    <G P, R> = <P, G^T R> between the batched forward and adjoint per model."""
    import torch
    w = W.ensemble_256(n=n, nz=nz)
    Pt = torch.from_numpy(w.rho).to(dev)
    R = residuals(13, 256, w.n_stations)
    Rt = torch.from_numpy(R).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        X = p.adjoint_batched(Rt).cpu().numpy()
        X2 = p.adjoint_batched(Rt).cpu().numpy()
        Y = p.forward_batched(Pt).cpu().numpy()
        p.set_batched_tuning(split=2)
        Xh = p.adjoint_batched(Rt).cpu().numpy()
        p.set_batched_tuning(split=3)
        for b in (0, 85, 170, 255):
            a64 = p.adjoint(torch.from_numpy(R[b].astype(np.float64)).to(dev)).cpu().numpy()
            assert relerr(X[b], a64) < 1e-5, b
            assert relerr(Xh[b], a64) < 1e-4, b
    assert np.array_equal(X, X2)
    lhs = np.einsum("bm,bm->b", Y, R.astype(np.float64))
    rhs = np.einsum("bn,bn->b", w.rho.astype(np.float64), X.astype(np.float64))
    scale = np.einsum("bm,bm->b", np.abs(Y), np.abs(R.astype(np.float64)))
    assert np.max(np.abs(lhs - rhs) / scale) < 1e-5


def test_batched_adjoint_one_slice_ragged(dev, orc):
    """One K-slice stored from the accumulators with N % 4 != 0 (per-element
    stores) and a partial last cell tile: 101x53x13 cells, 300 stations (19
    chunks), 250 vectors (8 tiles of 32, the last partial); 4 vectors against
    the oracle."""
    import torch
    g = W.Grid(101, 53, 13, -500.0, 200.0, -40.0, 40.0, 45.0, 30.0)
    c = g.centres()
    st = np.stack([W.uniform(31, a, 300, c[:, a].min(), c[:, a].max()) for a in range(2)]
                  + [W.uniform(31, 2, 300, 350.0, 450.0)], axis=1)
    R = residuals(32, 250, 300)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
    for b in (0, 100, 224, 249):
        want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
        assert relerr(X[b], want) < 1e-5, b


def test_batched_adjoint_rejects_and_empty(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=gfm.PRISM) as p:
        with pytest.raises(gfm.ZtmError):
            p.adjoint_batched(torch.zeros(2, w.n_stations, device=dev))
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        with pytest.raises(gfm.ZtmError):
            p.adjoint_batched(torch.zeros(257, w.n_stations, device=dev))
        out = p.adjoint_batched(torch.zeros(0, w.n_stations, device=dev))
        assert out.shape == (0, w.grid.n_cells)
        X = p.adjoint_batched(torch.zeros(3, w.n_stations, device=dev)).cpu().numpy()
        assert not np.any(X)
    # no stations: G^T R is zero (the output is cleared, not left as it was)
    with gfm.Plan(w.grid, np.zeros((0, 3)), kappa=w.kappa) as p:
        out = torch.full((5, w.grid.n_cells), 7.0, device=dev)
        p.adjoint_batched(torch.zeros(5, 0, device=dev), out=out)
        assert not torch.any(out).item()


def test_batched_forward_adjoint_full_size(dev, orc):
    """ensemble_256 at its full size (96x96x48 cells of 1 km, 9,216 stations,
    256 models and 256 residual vectors), default split: forward and adjoint
    of models / vectors 0, 85, 170 and 255 against the FP64 oracle on a seeded
    sample (64 stations over all cells; 2,048 cells over all stations) and
    against the FP64 single-model kernels on 0 and 255 everywhere (product bar
    1e-5 of max |.|), and <G P, R> = <P, G^T R> for every model
    (size-independent; 1e-5 of sum |G P| |R|)."""
    import torch
    w = W.ensemble_256()
    R = residuals(14, 256, w.n_stations)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        Pt = torch.from_numpy(w.rho).to(dev)
        Y = p.forward_batched(Pt).cpu().numpy()
        del Pt
        X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
        for b in (0, 255):
            g64 = p.forward(torch.from_numpy(w.rho[b].astype(np.float64)).to(dev)).cpu().numpy()
            a64 = p.adjoint(torch.from_numpy(R[b].astype(np.float64)).to(dev)).cpu().numpy()
            assert relerr(Y[b], g64) < 1e-5, b
            assert relerr(X[b], a64) < 1e-5, b
    st = np.unique((W.uniform(15, 1, 64) * w.n_stations).astype(np.int64))
    cells = np.unique((W.uniform(15, 2, 2048) * w.grid.n_cells).astype(np.int64))
    for b in (0, 85, 170, 255):
        g_or = orc.pm_forward(w.grid, w.kappa, w.rho[b].astype(np.float64), w.stations[st])
        a_or = orc.pm_adjoint_cells(w.grid, w.kappa, w.stations, R[b].astype(np.float64), cells)
        assert relerr(Y[b][st], g_or) < 1e-5, b
        assert relerr(X[b][cells], a_or) < 1e-5, b
    R64 = R.astype(np.float64)
    lhs = np.einsum("bm,bm->b", Y, R64)
    rhs = np.einsum("bn,bn->b", w.rho.astype(np.float64), X.astype(np.float64))
    scale = np.einsum("bm,bm->b", np.abs(Y), np.abs(R64))
    assert np.max(np.abs(lhs - rhs) / scale) < 1e-5


def test_small_ensembles_take_the_multi_field_forward(dev, orc):
    """1-4 models go through the multi-field FP32 forward (pair terms formed
    once, the plan's FP64 near-field correction per model): each model equals
    the single-model FP32 forward up to FP64 rounding of the partial sums,
    and the oracle to 1e-6; on a ragged grid with a threshold and stations
    inside the grid's footprint, close to the top face."""
    import torch
    g = W.Grid(37, 29, 11, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0)
    c = g.centres()
    st = np.stack([W.uniform(41, a, 300, c[:, a].min(), c[:, a].max()) for a in range(2)]
                  + [W.uniform(41, 2, 300, 50.0, 80.0)], axis=1)
    for B in (1, 2, 3, 4):
        P = models(42 + B, B, g.n_cells)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI, threshold=5.0) as p:
            Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
            singles = [p.forward(torch.from_numpy(P[b]).to(dev)).cpu().numpy() for b in range(B)]
        for b in range(B):
            assert relerr(Y[b], singles[b]) < 1e-12, (B, b)
            want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st, threshold=5.0)
            assert relerr(Y[b], want) < 1e-6, (B, b)


def test_small_batched_adjoint_is_the_fp32_adjoint(dev):
    """1-4 residual vectors run the FP32 adjoint per vector (the transpose of
    the batched forward's FP32 pass for up to 4 models): bitwise equal to
    ztm_plan_adjoint, and <G P, R> = <P, G^T R> to FP32 rounding."""
    import torch
    w = W.pm_cpu_small()
    for B in (1, 3, 4):
        R = residuals(50 + B, B, w.n_stations)
        P = models(60 + B, B, w.grid.n_cells)
        with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
            X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
            Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
            singles = [p.adjoint(torch.from_numpy(R[b]).to(dev)).cpu().numpy() for b in range(B)]
        for b in range(B):
            assert np.array_equal(X[b], singles[b]), (B, b)
            lhs = float(np.dot(Y[b], R[b].astype(np.float64)))
            rhs = float(np.dot(P[b].astype(np.float64), X[b].astype(np.float64)))
            scale = float(np.dot(np.abs(Y[b]), np.abs(R[b].astype(np.float64))))
            assert abs(lhs - rhs) < 1e-6 * scale, (B, b)


def test_batched_random_geometries(dev, orc):
    """Seeded sweep over ragged grids (signed spacings, nx from 3 to 70),
    stations above the top face, 1-40 models / vectors: the batched forward
    (MFMA or, for up to 4 models, the multi-field FP32 pass) and the batched
    adjoint against the oracle on the first and last model / vector."""
    import torch
    rng = np.random.default_rng(20190317)
    for case in range(6):
        nx, ny, nz = (int(v) for v in rng.integers((3, 2, 2), (71, 40, 12)))
        d = rng.uniform(20.0, 80.0, 3) * rng.choice((-1.0, 1.0), 3)
        g = W.Grid(nx, ny, nz, float(rng.uniform(-1e3, 1e3)), float(rng.uniform(-1e3, 1e3)),
                   0.0, float(d[0]), float(d[1]), float(d[2]))
        c = g.centres()
        m = int(rng.integers(1, 300))
        top = c[:, 2].max() + abs(d[2]) / 2
        st = np.stack([rng.uniform(c[:, 0].min(), c[:, 0].max(), m),
                       rng.uniform(c[:, 1].min(), c[:, 1].max(), m),
                       rng.uniform(top + 5.0, top + 100.0, m)], axis=1)
        B = int(rng.choice((1, 3, 4, 5, 17, 40)))
        P = rng.uniform(-300, 300, (B, g.n_cells)).astype(np.float32)
        R = rng.normal(0.0, 0.5, (B, m)).astype(np.float32)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
            X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
        for b in sorted({0, B - 1}):
            want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st)
            assert relerr(Y[b], want) < 1e-5, (case, nx, ny, nz, m, B, b)
            want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
            assert relerr(X[b], want) < 1e-5, (case, nx, ny, nz, m, B, b)


@pytest.mark.parametrize("n,nz", [(48, 24), (64, 12)])
def test_batched_fp16_split_half_size(dev, n, nz):
    """Split code 4 (FP16 hi/lo parts under power-of-two scales: G by the
    plan's bound on |G|, each model / vector by its largest entry): forward
    and adjoint of 256 models against the FP64 single-model kernels on 4 of
    them, bitwise repeatability and the dot-product identity."""
    import torch
    w = W.ensemble_256(n=n, nz=nz)
    Pt = torch.from_numpy(w.rho).to(dev)
    R = residuals(13, 256, w.n_stations)
    Rt = torch.from_numpy(R).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_batched_tuning(split=4)
        Y = p.forward_batched(Pt).cpu().numpy()
        X = p.adjoint_batched(Rt).cpu().numpy()
        assert np.array_equal(Y, p.forward_batched(Pt).cpu().numpy())
        assert np.array_equal(X, p.adjoint_batched(Rt).cpu().numpy())
        for b in (0, 85, 170, 255):
            g64 = p.forward(torch.from_numpy(w.rho[b].astype(np.float64)).to(dev)).cpu().numpy()
            a64 = p.adjoint(torch.from_numpy(R[b].astype(np.float64)).to(dev)).cpu().numpy()
            assert relerr(Y[b], g64) < 5e-6, b
            assert relerr(X[b], a64) < 5e-6, b
    lhs = np.einsum("bm,bm->b", Y, R.astype(np.float64))
    rhs = np.einsum("bn,bn->b", w.rho.astype(np.float64), X.astype(np.float64))
    scale = np.einsum("bm,bm->b", np.abs(Y), np.abs(R.astype(np.float64)))
    assert np.max(np.abs(lhs - rhs) / scale) < 1e-5


def test_batched_fp16_split_random_geometries(dev, orc):
    """Split code 4 over seeded ragged grids whose stations range from a
    hundredth of a cell above the top face (a large bound on |G|) to far
    above it (a small one), models of very different magnitudes and a zero
    model: forward and adjoint against the oracle."""
    import torch
    rng = np.random.default_rng(6021023)
    for case in range(6):
        nx, ny, nz = (int(v) for v in rng.integers((3, 2, 2), (71, 40, 12)))
        d = rng.uniform(20.0, 80.0, 3) * rng.choice((-1.0, 1.0), 3)
        g = W.Grid(nx, ny, nz, 0.0, 0.0, 0.0, float(d[0]), float(d[1]), float(d[2]))
        c = g.centres()
        m = int(rng.integers(5, 300))
        top = c[:, 2].max() + abs(d[2]) / 2
        lift = abs(d[2]) * (0.01 if case % 2 == 0 else 20.0)
        st = np.stack([rng.uniform(c[:, 0].min(), c[:, 0].max(), m),
                       rng.uniform(c[:, 1].min(), c[:, 1].max(), m),
                       rng.uniform(top + lift, top + 2 * lift, m)], axis=1)
        B = int(rng.choice((9, 33, 64)))
        mag = 10.0 ** rng.uniform(-6, 6, (B, 1))
        P = (rng.uniform(-1, 1, (B, g.n_cells)) * mag).astype(np.float32)
        P[B // 2] = 0.0
        R = (rng.normal(0.0, 1.0, (B, m)) * mag).astype(np.float32)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            p.set_batched_tuning(split=4)
            Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
            X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
        assert not Y[B // 2].any()
        for b in sorted({0, B - 1}):
            want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st)
            assert relerr(Y[b], want) < 1e-5, (case, nx, ny, nz, m, B, b)
            want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
            assert relerr(X[b], want) < 1e-5, (case, nx, ny, nz, m, B, b)


def test_batched_fp16_split_adjoint_split_k(dev, orc):
    """Split code 4 with more station chunks than one K-slice takes (13,000
    stations over 4,096 cells): the adjoint's slice partials go through the
    reduce with the per-vector scales, against the oracle and the BF16
    three-way split."""
    import torch
    g = W.Grid(32, 16, 8, 0.0, 0.0, 0.0, 50.0, 50.0, 40.0)
    c = g.centres()
    rng = np.random.default_rng(77)
    m = 13000
    st = np.stack([rng.uniform(c[:, 0].min(), c[:, 0].max(), m),
                   rng.uniform(c[:, 1].min(), c[:, 1].max(), m),
                   rng.uniform(c[:, 2].max() + 25.0, c[:, 2].max() + 30.0, m)], axis=1)
    R = (rng.normal(0.0, 1.0, (40, m)) * 10.0 ** rng.uniform(-3, 3, (40, 1))).astype(np.float32)
    Rt = torch.from_numpy(R).to(dev)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        p.set_batched_tuning(split=3)
        X3 = p.adjoint_batched(Rt).cpu().numpy()
        p.set_batched_tuning(split=4)
        assert p.batched_tuning()[2] == 4
        X = p.adjoint_batched(Rt).cpu().numpy()
    for b in (0, 21, 39):
        want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
        assert relerr(X[b], want) < 1e-5, b
        assert relerr(X3[b], want) < 1e-5, b


def test_batched_fp16_split_edge_plans(dev):
    """Split code 4 where the bound on |G| is 0: no stations (the adjoint
    writes zeros), and a one-layer grid with every station in the layer's
    plane (G is 0 throughout, so are both directions)."""
    import torch
    g = W.Grid(20, 10, 1, 0.0, 0.0, 0.0, 50.0, 50.0, 40.0)
    P = torch.ones(6, g.n_cells, device=dev)
    with gfm.Plan(g, np.zeros((0, 3)), kappa=W.KAPPA_SI) as p:
        p.set_batched_tuning(split=4)
        X = p.adjoint_batched(torch.ones(6, 0, device=dev)).cpu().numpy()
        assert X.shape == (6, g.n_cells) and not X.any()
    c = g.centres()
    st = np.stack([np.linspace(c[:, 0].min(), c[:, 0].max(), 7) + 3.0,
                   np.linspace(c[:, 1].min(), c[:, 1].max(), 7) - 2.0,
                   np.full(7, c[0, 2])], axis=1)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        p.set_batched_tuning(split=4)
        Y = p.forward_batched(P).cpu().numpy()
        X = p.adjoint_batched(torch.ones(6, 7, device=dev)).cpu().numpy()
    assert np.isfinite(Y).all() and not Y.any()
    assert np.isfinite(X).all() and not X.any()


def test_batched_fp16_split_mixed_heights(dev, orc):
    """Split code 4 with stations 0.05 cell above the grid's top face and
    stations 2,000 cells above it (their largest |G| differ by ~2^23.6): the
    forward's per-station G scales keep the far stations' results to 1e-5 of
    their own largest value (an FP16 emulation with one plan-wide scale gives
    1.3e-4 there, per-station scales 2.0e-6); the adjoint, whose one G scale
    cannot span that range, runs the BF16 three-way split on this plan, also
    with residuals on the far stations only."""
    import torch
    g = W.Grid(40, 40, 10, 0.0, 0.0, 0.0, 50.0, 50.0, 50.0)
    c = g.centres()
    rng = np.random.default_rng(5)
    m = 160
    top = c[:, 2].max() + 25.0
    z = np.where(np.arange(m) % 2 == 0, top + 2.5, top + 100000.0)
    st = np.stack([rng.uniform(c[:, 0].min(), c[:, 0].max(), m),
                   rng.uniform(c[:, 1].min(), c[:, 1].max(), m), z], axis=1)
    far = np.arange(m) % 2 == 1
    P = rng.uniform(-300, 300, (8, g.n_cells)).astype(np.float32)
    Rv = rng.normal(0.0, 1.0, (8, m)).astype(np.float32)
    Rf = Rv * far
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        p.set_batched_tuning(split=4)
        assert p.batched_tuning() == (0, 4, 3)
        Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
        X = p.adjoint_batched(torch.from_numpy(Rv).to(dev)).cpu().numpy()
        Xf = p.adjoint_batched(torch.from_numpy(Rf).to(dev)).cpu().numpy()
    for b in (0, 7):
        want = orc.pm_forward(g, W.KAPPA_SI, P[b].astype(np.float64), st)
        assert relerr(Y[b], want) < 1e-5, b
        assert relerr(Y[b][far], want[far]) < 1e-5, b
        want = orc.pm_adjoint(g, W.KAPPA_SI, st, Rv[b].astype(np.float64))
        assert relerr(X[b], want) < 1e-5, b
        want = orc.pm_adjoint(g, W.KAPPA_SI, st, Rf[b].astype(np.float64))
        assert relerr(Xf[b], want) < 5e-6, b


def test_batched_fp16_adjoint_far_residuals(dev, orc):
    """Residuals on far stations only. Stations at one height over the grid
    (bounds on |G| within a factor 3): the FP16 adjoint stays on and meets
    5e-6 of max |G^T r| against the oracle with residuals only in one corner
    of the station grid (the split itself adds ~1e-7 in an FP16 emulation;
    the rest is the MFMA's FP32 accumulation). The same stations plus one
    10^-3 cell from a cell centre (its bound ~10^5 times the others', outside
    the FP16 adjoint's range): the adjoint runs the BF16 three-way split and
    meets the oracle with the near station's residual 0."""
    import torch
    g = W.Grid(40, 40, 10, 0.0, 0.0, -25.0, 50.0, 50.0, -50.0)
    c = g.centres()
    xs = np.linspace(c[:, 0].min(), c[:, 0].max(), 30)
    X0, Y0 = np.meshgrid(xs, xs)
    st = np.stack([X0.ravel(), Y0.ravel(), np.full(X0.size, 10.0)], axis=1)
    corner = (st[:, 0] > c[:, 0].max() - 300.0) & (st[:, 1] > c[:, 1].max() - 300.0)
    rng = np.random.default_rng(12)
    R = (rng.normal(0.0, 1.0, (40, st.shape[0])) * corner).astype(np.float32)
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        assert p.batched_tuning() == (0, 4, 4)
        X = p.adjoint_batched(torch.from_numpy(R).to(dev)).cpu().numpy()
    for b in (0, 39):
        want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].astype(np.float64))
        assert relerr(X[b], want) < 5e-6, b
    near = np.array([[c[1000, 0] + 0.05, c[1000, 1], c[1000, 2] + 0.02]])
    st2 = np.concatenate([st, near])
    R2 = np.concatenate([R, np.zeros((40, 1), np.float32)], axis=1)
    with gfm.Plan(g, st2, kappa=W.KAPPA_SI) as p:
        assert p.batched_tuning() == (0, 4, 3)
        X2 = p.adjoint_batched(torch.from_numpy(R2).to(dev)).cpu().numpy()
    for b in (0, 39):
        want = orc.pm_adjoint(g, W.KAPPA_SI, st2, R2[b].astype(np.float64))
        assert relerr(X2[b], want) < 5e-6, b


def test_batched_fp16_tiny_models(dev, orc):
    """Split code 4 on models whose largest entry is ~1e-36 (the power-of-two
    model scale is held to 2^120, so the parts stay finite) beside ordinary
    ones: every model against the oracle."""
    import torch
    w = W.pm_cpu_small()
    P = models(19, 40, w.grid.n_cells).astype(np.float64)
    P[3] *= 1e-38
    P[17] *= 1e-30
    P = P.astype(np.float32)
    assert P[3].any()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_batched_tuning(split=4)
        Y = p.forward_batched(torch.from_numpy(P).to(dev)).cpu().numpy()
    assert np.isfinite(Y).all()
    for b in (0, 3, 17):
        want = orc.pm_forward(w.grid, w.kappa, P[b].astype(np.float64), w.stations)
        assert relerr(Y[b], want) < 1e-5, b


@pytest.mark.parametrize("split", [4, 2])
def test_batched_kernel_variants_agree(dev, orc, split):
    """Kernel variants 0, 1 and 2 of the two-part splits issue the same MFMA
    products in the same order per accumulator: the forward is bitwise
    equal, on grids that take each cell-position mode (nx < 8, nx >= 8 not a
    multiple of 8, a multiple of 8) and with one and several K-slices; the
    adjoint is bitwise equal through the slice reduce and, with one slice,
    within 1e-7 of max |X| where variant 1 applies FP16's per-vector factors in
    its epilogue (float) instead of the transpose (double). Both against the
    oracle."""
    import torch
    for nx, ny, nz, m in ((5, 7, 9, 40), (37, 5, 6, 130), (64, 8, 4, 300)):
        g = W.Grid(nx, ny, nz, 10.0, 20.0, -15.0, 30.0, 35.0, -25.0)
        c = g.centres()
        st = np.stack([W.uniform(8, a, m, c[:, a].min() - 50, c[:, a].max() + 50)
                       for a in range(2)] + [np.full(m, 40.0)], axis=1)
        P = torch.from_numpy(models(9 + nx, 40, g.n_cells)).to(dev)
        R = torch.from_numpy(residuals(10 + nx, 40, m)).to(dev)
        with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
            for ks in (0, 3):
                out = {}
                for var in (0, 1, 2):
                    p.set_batched_tuning(slices=ks, split=split)
                    p.set_batched_variant(var)
                    assert p.batched_settings()[3] == var
                    out[var] = (p.forward_batched(P).cpu().numpy(),
                                p.adjoint_batched(R).cpu().numpy())
                assert np.array_equal(out[0][0], out[1][0]), (nx, ks)
                assert np.array_equal(out[0][0], out[2][0]), (nx, ks)
                assert np.array_equal(out[1][1], out[2][1]), (nx, ks)
                if ks == 3 or split == 2:
                    assert np.array_equal(out[0][1], out[1][1]), (nx, ks)
                else:
                    assert relerr(out[1][1], out[0][1].astype(np.float64)) < 1e-7, (nx, ks)
                tol = 1e-5 if split == 4 else 1e-4
                for b in (0, 39):
                    want = orc.pm_forward(g, W.KAPPA_SI, P[b].cpu().numpy().astype(np.float64), st)
                    assert relerr(out[1][0][b], want) < tol, (nx, ks, b)
                    want = orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].cpu().numpy().astype(np.float64))
                    assert relerr(out[1][1][b], want) < tol, (nx, ks, b)


@pytest.mark.parametrize("split", [4, 2])
def test_batched_autotune_candidates_and_choice(dev, orc, split):
    """Every kernel setting the autotuner chooses from (variant 0; variants 1
    and 2 with 4 and 8 waves per workgroup, i.e. 128 or 256 stations / cells
    per workgroup; automatic and explicit K-slices) gives the oracle's forward and
    adjoint to 1e-5 (BF16 2-way: 1e-4) on a ragged grid with 40 models. After
    autotune_batched each direction's setting is one of the candidates, and
    the calls run it: their results equal, bit for bit, those of the same
    setting set by hand on a fresh plan."""
    import torch
    g = W.Grid(37, 29, 11, 100.0, -50.0, 30.0, -70.0, 45.0, -25.0)
    c = g.centres()
    m = 700
    st = np.stack([W.uniform(27, a, m, c[:, a].min() - 100, c[:, a].max() + 100)
                   for a in range(2)] + [np.full(m, 90.0)], axis=1)
    P = torch.from_numpy(models(28, 40, g.n_cells)).to(dev)
    R = torch.from_numpy(residuals(29, 40, m)).to(dev)
    tol = 1e-5 if split == 4 else 1e-4
    want_f = [orc.pm_forward(g, W.KAPPA_SI, P[b].cpu().numpy().astype(np.float64), st)
              for b in (0, 39)]
    want_a = [orc.pm_adjoint(g, W.KAPPA_SI, st, R[b].cpu().numpy().astype(np.float64))
              for b in (0, 39)]
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as p:
        p.set_batched_tuning(split=split)
        for var, waves in ((0, 4), (1, 4), (1, 8), (2, 4), (2, 8)):
            for ks in (0, 2, 7):
                for d in (0, 1):
                    p.set_batched_kernel(d, 40, var, waves, ks)
                    assert p.batched_kernel(d, 40) == (var, waves, ks)
                Y = p.forward_batched(P).cpu().numpy()
                X = p.adjoint_batched(R).cpu().numpy()
                for q, b in enumerate((0, 39)):
                    assert relerr(Y[b], want_f[q]) < tol, (var, waves, ks, b)
                    assert relerr(X[b], want_a[q]) < tol, (var, waves, ks, b)
        p.autotune_batched(40, reps=2)
        kf, ka = p.batched_kernel(0, 40), p.batched_kernel(1, 40)
        assert (kf[0], kf[1]) in ((0, 4), (1, 4), (1, 8), (2, 4), (2, 8)) and kf[2] >= 1, kf
        assert (ka[0], ka[1]) in ((0, 4), (1, 4), (1, 8)) and ka[2] >= 1, ka
        Y = p.forward_batched(P).cpu().numpy()
        X = p.adjoint_batched(R).cpu().numpy()
    with gfm.Plan(g, st, kappa=W.KAPPA_SI) as q:
        q.set_batched_tuning(split=split)
        q.set_batched_kernel(0, 40, *kf)
        q.set_batched_kernel(1, 40, *ka)
        assert np.array_equal(q.forward_batched(P).cpu().numpy(), Y)
        assert np.array_equal(q.adjoint_batched(R).cpu().numpy(), X)
    for q_, b in enumerate((0, 39)):
        assert relerr(Y[b], want_f[q_]) < tol and relerr(X[b], want_a[q_]) < tol
