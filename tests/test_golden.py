"""Committed golden vectors (tests/golden/, written by make_golden.py).

CPU: the seeded generators still produce the recorded inputs, and the oracle
still produces the recorded outputs. GPU: the HIP path matches them.
"""
import json
import os

import numpy as np
import pytest

from ztm_amd import gfm, workloads as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(HERE, "manifest.json")))


def load(name):
    return np.load(os.path.join(HERE, MAN[name]["file"]), allow_pickle=False)


def check_sum(a, rec):
    a = np.asarray(a, dtype=np.float64).ravel()
    assert a.size == rec["n"]
    assert a[0] == rec["first"] and a[-1] == rec["last"]
    assert abs(a.sum() - rec["sum"]) <= 1e-9 * max(1.0, abs(rec["sumsq"]) ** 0.5)
    assert abs((a * a).sum() - rec["sumsq"]) <= 1e-12 * rec["sumsq"]


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


# ------------------------------------------------------------------- CPU
def test_generators_reproduce_recorded_inputs():
    w = W.pm_cpu_small()
    check_sum(w.rho, MAN["pm_cpu_small"]["rho"])
    check_sum(w.stations, MAN["pm_cpu_small"]["stations"])
    rho, s = W.tfm_shaped()
    check_sum(rho, MAN["tfm_shaped"]["rho"])
    check_sum(s, MAN["tfm_shaped"]["s"])
    check_sum(W.adjoint_small().extra["residual"], MAN["adjoint_small"]["r"])


def test_oracle_reproduces_golden(orc):
    w = W.pm_cpu_small()
    assert relerr(orc.pm_forward(w.grid, w.kappa, w.rho, w.stations),
                  load("pm_cpu_small")) < 1e-13
    rho, s = W.tfm_shaped()
    assert relerr(orc.gf(10, 61, 61, 33, 1, 2, rho, s), load("tfm_shaped")) < 1e-13
    w = W.adjoint_small()
    assert relerr(orc.pm_adjoint(w.grid, w.kappa, w.stations, w.extra["residual"]),
                  load("adjoint_small")) < 1e-13


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_matches_golden(dev):
    import torch
    w = W.pm_cpu_small()
    with gfm.Plan(w.grid, w.stations, kappa=w.kappa) as p:
        g64 = p.forward(torch.from_numpy(w.rho).to(dev)).cpu().numpy()
        g32 = p.forward(torch.from_numpy(w.rho.astype(np.float32)).to(dev)).cpu().numpy()
        r = W.adjoint_small().extra["residual"]
        a64 = p.adjoint(torch.from_numpy(r).to(dev)).cpu().numpy()
        a32 = p.adjoint(torch.from_numpy(r.astype(np.float32)).to(dev)).cpu().numpy()
    assert relerr(g64, load("pm_cpu_small")) < 1e-12
    assert relerr(g32, load("pm_cpu_small")) < 1e-5
    assert relerr(a64, load("adjoint_small")) < 1e-12
    assert relerr(a32, load("adjoint_small")) < 1e-5
    rho, s = W.tfm_shaped()
    g = gfm.gfm_host(10, 61, 61, 33, 1, 2, rho, s)
    assert relerr(g, load("tfm_shaped")) < 1e-6
