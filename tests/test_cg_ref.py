"""Pins of the CPU CG checker used for the batched solver: L^T L against an
explicitly assembled difference matrix, and CG against a direct solve."""
import numpy as np

from cg_ref import cg_normal, ltl


def difference_matrix(nx, ny, nz):
    """One row per interior edge along i, j and k: +1 at the far cell, -1 at
    the near one."""
    rows = []
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    for ax in range(3):
        a = np.moveaxis(idx, ax, 0)
        for lo, hi in zip(a[:-1].ravel(), a[1:].ravel()):
            r = np.zeros(nx * ny * nz)
            r[hi], r[lo] = 1.0, -1.0
            rows.append(r)
    return np.array(rows)


def test_ltl_matches_assembled_operator():
    nx, ny, nz = 4, 3, 5
    L = difference_matrix(nx, ny, nz)
    assert L.shape[0] == (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    v = np.random.default_rng(0).normal(size=nx * ny * nz)
    np.testing.assert_allclose(ltl(v, nx, ny, nz), L.T @ (L @ v), rtol=0, atol=1e-12)
    assert np.allclose(ltl(np.ones(nx * ny * nz), nx, ny, nz), 0.0)
    # diagonal = number of axis neighbours (3 at a corner, 6 inside)
    diag = np.diag(L.T @ L).reshape(nz, ny, nx)
    assert diag[0, 0, 0] == 3 and diag[2, 1, 1] == 6


def test_cg_reaches_the_direct_solution():
    rng = np.random.default_rng(1)
    nx, ny, nz = 5, 4, 3
    n, m, lam = nx * ny * nz, 25, 0.3
    G = rng.normal(size=(m, n))
    d = rng.normal(size=m)
    L = difference_matrix(nx, ny, nz)
    A = G.T @ G + lam * L.T @ L
    b = G.T @ d
    x, rr = cg_normal(lambda v: G @ v, lambda r: G.T @ r, lambda v: ltl(v, nx, ny, nz), d, lam,
                      200)
    want = np.linalg.solve(A, b)
    np.testing.assert_allclose(x, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())
    assert rr[0] == b @ b
    # after a few steps the recurrence residual equals the true one
    x5, rr5 = cg_normal(lambda v: G @ v, lambda r: G.T @ r, lambda v: ltl(v, nx, ny, nz), d,
                        lam, 5)
    true5 = b - A @ x5
    assert abs(rr5[-1] - true5 @ true5) <= 1e-9 * rr5[0]
    assert np.all(np.diff(rr[:20]) != 0)
