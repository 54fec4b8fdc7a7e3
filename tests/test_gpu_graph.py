"""Plan calls captured by the caller in a torch CUDA (HIP) graph: every replay,
with eager work between replays, equals the eager result bit for bit."""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("formula", [gfm.POINT_MASS, gfm.PRISM])
def test_plan_forward_adjoint_in_graph(dev, formula):
    import torch
    w = W.pm_cpu_small()
    rho = torch.from_numpy(w.rho.astype(np.float32)).to(dev)
    r = torch.from_numpy(W.normal(3, 7, w.n_stations, 0.5).astype(np.float32)).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa, formula=formula) as p:
        g = torch.empty(w.n_stations, dtype=torch.float64, device=dev)
        a = torch.empty(w.grid.n_cells, dtype=torch.float32, device=dev)
        p.forward(rho, out=g)                 # grows the workspace before capture
        p.adjoint(r, out=a)
        torch.cuda.synchronize()
        g0, a0 = g.clone(), a.clone()
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph):
                p.forward(rho, out=g)
                p.adjoint(r, out=a)
        torch.cuda.current_stream().wait_stream(side)
        for rep in range(3):
            g.fill_(-1.0)
            a.fill_(-1.0)
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(g, g0), rep
            assert torch.equal(a, a0), rep
            p.forward(2.0 * rho)              # eager work between replays
            torch.cuda.synchronize()


@pytest.mark.parametrize("split", [3, 4])
def test_batched_forward_in_graph(dev, split):
    """Batched forward (split 3: pack, MFMA kernel, reduce; split 4 adds the
    scale clear and the absmax with its atomics) captured and replayed 3x
    with a smaller eager call between replays, bitwise equal to the eager
    result."""
    import torch
    w = W.pm_cpu_small()
    P = torch.from_numpy(np.stack([W.uniform(5, b, w.grid.n_cells, -300, 300)
                                   for b in range(40)]).astype(np.float32)).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_batched_tuning(split=split)
        Y = torch.empty(40, w.n_stations, dtype=torch.float64, device=dev)
        p.forward_batched(P, out=Y)
        torch.cuda.synchronize()
        Y0 = Y.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            p.forward_batched(P, out=Y)
        for rep in range(3):
            Y.zero_()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(Y, Y0), rep
            p.forward_batched(P[:8].contiguous())
            torch.cuda.synchronize()


@pytest.mark.parametrize("split", [3, 4])
def test_batched_adjoint_in_graph(dev, split):
    """Batched adjoint (pack, MFMA kernel, reduce: no memset; split 4 adds
    the scale clear, the absmax and the transpose) captured and replayed 3x
    with a smaller eager call between replays, bitwise equal to the eager
    result."""
    import torch
    w = W.pm_cpu_small()
    R = torch.from_numpy(np.stack([W.normal(6, b, w.n_stations, 0.5)
                                   for b in range(40)]).astype(np.float32)).to(dev)
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        p.set_batched_tuning(split=split)
        X = torch.empty(40, w.grid.n_cells, dtype=torch.float32, device=dev)
        p.adjoint_batched(R, out=X)
        torch.cuda.synchronize()
        X0 = X.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            p.adjoint_batched(R, out=X)
        for rep in range(3):
            X.zero_()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(X, X0), rep
            p.adjoint_batched(R[:8].contiguous())
            torch.cuda.synchronize()


def test_zero_station_adjoints_in_graph(dev):
    """Adjoints over a plan with no stations clear their output with a kernel
    (no memset), so they capture: single-vector FP32 and FP64, the batched
    adjoint with 3 vectors (the per-vector FP32 path) and with 40, and the
    stored adjoint; every replay rewrites the zeros over a poisoned output."""
    import torch
    g = W.Grid(9, 7, 5, 0.0, 0.0, -500.0, 100.0, 100.0, 100.0)
    with gfm.Plan(g, np.zeros((0, 3))) as p:
        r32 = torch.zeros(0, dtype=torch.float32, device=dev)
        r64 = torch.zeros(0, dtype=torch.float64, device=dev)
        R3 = torch.zeros(3, 0, dtype=torch.float32, device=dev)
        R40 = torch.zeros(40, 0, dtype=torch.float32, device=dev)
        a32 = torch.full((g.n_cells,), 7.0, dtype=torch.float32, device=dev)
        a64 = torch.full((g.n_cells,), 7.0, dtype=torch.float64, device=dev)
        X3 = torch.full((3, g.n_cells), 7.0, dtype=torch.float32, device=dev)
        X40 = torch.full((40, g.n_cells), 7.0, dtype=torch.float32, device=dev)
        s32 = torch.full((g.n_cells,), 7.0, dtype=torch.float32, device=dev)
        p.store()
        outs = (a32, a64, X3, X40, s32)

        def calls():
            p.adjoint(r32, out=a32)
            p.adjoint(r64, out=a64)
            p.adjoint_batched(R3, out=X3)
            p.adjoint_batched(R40, out=X40)
            p.adjoint_stored(r32, out=s32)

        calls()
        torch.cuda.synchronize()
        assert all(not o.any() for o in outs)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            calls()
        for rep in range(3):
            for o in outs:
                o.fill_(float(rep + 1))
            graph.replay()
            torch.cuda.synchronize()
            assert all(not o.any() for o in outs), rep


def test_gauss_newton_la_in_graph(dev):
    """The dense LA of a Gauss-Newton step (normal equations, the 12 damped
    steps, the inverse diagonal) on caller buffers, captured and replayed 3x
    with eager calls on other buffers between replays: every replay equals
    the eager result bit for bit (no call allocates or issues a memset)."""
    import torch
    rows, n = 2000, 100
    J = torch.from_numpy(W.normal(11, 1, rows * n).reshape(rows, n)).to(dev)
    r = torch.from_numpy(W.normal(11, 2, rows)).to(dev)
    sig = torch.from_numpy(W.uniform(11, 3, n, 0.05, 4.0)).to(dev)
    b = torch.from_numpy(W.normal(11, 4, n)).to(dev)
    dm = torch.from_numpy(W.GN_DAMPING).to(dev)
    nd = dm.shape[0]
    f64 = dict(dtype=torch.float64, device=dev)
    JJ, jtr, B, V = (torch.empty(n, n, **f64), torch.empty(n, **f64), torch.empty(nd, n, **f64),
                     torch.empty(n, **f64))
    work = torch.empty(gfm.gn_workspace_bytes(rows, n, nd), dtype=torch.uint8, device=dev)

    def calls():
        gfm.gn_normal(J, r, out=(JJ, jtr), work=work)
        gfm.gn_damped_steps(JJ, jtr, sig, dm, b, out=B, work=work)
        gfm.gn_inverse_diag(JJ, 1e-3, out=V, work=work)

    calls()
    torch.cuda.synchronize()
    want = [t.clone() for t in (JJ, jtr, B, V)]
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph):
            calls()
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        for t in (JJ, jtr, B, V):
            t.fill_(-1.0)
        graph.replay()
        torch.cuda.synchronize()
        for got, ref in zip((JJ, jtr, B, V), want):
            assert torch.equal(got, ref), rep
        JJ2, jtr2 = gfm.gn_normal(2.0 * J, r)          # eager work between replays
        gfm.gn_damped_steps(JJ2, jtr2, sig, dm, b)
        torch.cuda.synchronize()
