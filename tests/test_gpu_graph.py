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
