"""ztm_gf_noalloc (the reference-kernel call on a fixed per-GPU scratch array,
no allocation) against ztm_gf_async through the C ABI: bit for bit on two
calls in a row, the second on a smaller geometry, and a third back on the
first (the array keeps the last call's contents); both against the FP64
oracle; replays of a captured HIP graph of the call; and the size limit
rejected before anything runs."""
import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

pytestmark = pytest.mark.gpu


def _call(name, geom, g, rho, s):
    import torch
    di, dj, dk, fh, fv, bfr = geom
    stream = torch.cuda.current_stream().cuda_stream
    return getattr(gfm.lib(), name)(bfr, di, dj, dk, fh, fv, g.data_ptr(), rho.data_ptr(),
                                    s.data_ptr(), stream)


def _case(geom, seed):
    import torch
    di, dj, dk, fh, fv, bfr = geom
    rho, s = W.tfm_shaped(seed, di, dj, dk, fh, fv)
    g0 = W.uniform(seed, 9, di * dj, -1, 1).astype(np.float32)
    return (rho, s, g0, torch.from_numpy(rho).cuda(), torch.from_numpy(s).cuda(),
            torch.from_numpy(g0).cuda())


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_equals_gf_async(dev, orc, order):
    import torch
    big, small = (61, 61, 33, 1, 2, 10), (23, 17, 6, 1, 2, 3)
    geoms = [big, small, big] if order == 0 else [(12, 9, 4, 2, 3, 2), small]
    for c, geom in enumerate(geoms):
        print(f"[noalloc] call {c}: geom {geom}", flush=True)
        rho, s, g0, rt, st, gt = _case(geom, 40 + c)
        g_ref = gt.clone()
        g_new = gt.clone()
        assert _call("ztm_gf_async", geom, g_ref, rt, st) == 0
        assert _call("ztm_gf_noalloc", geom, g_new, rt, st) == 0, gfm.lib().ztm_last_error()
        torch.cuda.synchronize()
        assert torch.equal(g_new, g_ref), geom
        di, dj, dk, fh, fv, bfr = geom
        want = orc.gf(bfr, di, dj, dk, fh, fv, rho, s, g=g0.astype(np.float64))
        got = g_new.cpu().numpy()
        d = want - g0
        assert np.max(np.abs(got - want)) <= 1e-6 * max(np.max(np.abs(d)), 1e-30), geom


def test_noalloc_in_graph(dev):
    """Captured once, replayed three times with other work on the scratch
    array between replays (an eager call on another geometry): every replay
    equals the eager result."""
    import torch
    geom, other = (61, 61, 33, 1, 2, 10), (23, 17, 6, 1, 2, 3)
    _, _, _, rt, st, gt = _case(geom, 50)
    _, _, _, rt2, st2, gt2 = _case(other, 51)
    want = gt.clone()
    assert _call("ztm_gf_noalloc", geom, want, rt, st) == 0
    torch.cuda.synchronize()
    g = gt.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        rc = _call("ztm_gf_noalloc", geom, g, rt, st)
    assert rc == 0, gfm.lib().ztm_last_error()
    for rep in range(3):
        g.copy_(gt)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(g, want), rep
        assert _call("ztm_gf_noalloc", other, gt2.clone(), rt2, st2) == 0
        torch.cuda.synchronize()


def test_noalloc_rejects_oversized_geometry(dev):
    """250,000 interior stations need more than the 16 MiB array: ZTM_EINVAL
    and a message naming the limit, before any kernel runs (the buffers passed
    are never touched)."""
    import torch
    t = torch.zeros(16, dtype=torch.float32, device=dev)
    L = gfm.lib()
    rc = L.ztm_gf_noalloc(0, 500, 500, 2, 1, 1, t.data_ptr(), t.data_ptr(), t.data_ptr(),
                          torch.cuda.current_stream().cuda_stream)
    assert rc == 1
    msg = L.ztm_last_error().decode()
    assert "scratch" in msg and str(16 << 20) in msg, msg
    assert L.ztm_gf_noalloc(10, 61, 61, 33, 1, 2, None, t.data_ptr(), t.data_ptr(), None) == 1
