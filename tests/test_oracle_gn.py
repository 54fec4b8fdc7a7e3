"""Pins of the oracle's dense LA of the Gauss-Newton step (MTM, J^T r,
Gauss-Jordan inv, the damped steps of LMA::fit, diag(inv(J^T J + I d))):
against numpy's LAPACK routines, identities, pivoting and singular cases, and
-- where oracle/_ref was built from the reference's own zmath.cpp -- against
the reference's MTM and inv bit for bit."""
import numpy as np
import pytest

import oracle

# the reference's damping grid: dmp = {{0,.3,1,3,10,100},{1e4,1e5}} (TFM.cpp:202,209)
DM_REF = np.array([[a, b] for a in (0, .3, 1, 3, 10, 100) for b in (1e4, 1e5)])


def jacobian(seed, rows, cols):
    """A Jacobian shaped like the inversion's: columns of very different scales
    (event coordinates, velocities, Moho depths), some sparse."""
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(rows, cols)) * np.exp(rng.uniform(-3, 3, cols))
    J[rng.uniform(size=(rows, cols)) < 0.3] = 0.0
    return J


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_inv_matches_numpy_and_inverts():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 60):
        A = rng.normal(size=(n, n)) + n * np.eye(n) * rng.uniform(0.2, 1.0)
        ok, Ai = oracle.inv(A)
        assert ok
        assert relerr(Ai, np.linalg.inv(A)) < 1e-12
        assert np.max(np.abs(Ai @ A - np.eye(n))) < 1e-12


def test_inv_row_swaps():
    """Zero and tiny diagonal entries: the pivot search must swap rows."""
    P = np.eye(5)[[3, 0, 4, 1, 2]]
    ok, Pi = oracle.inv(P)
    assert ok and np.array_equal(Pi, P.T)
    A = np.array([[1e-14, 2.0, 0.0], [3.0, 1.0, 1.0], [0.0, 1.0, 4.0]])
    ok, Ai = oracle.inv(A)
    assert ok and np.max(np.abs(Ai @ A - np.eye(3))) < 1e-13


def test_inv_exact_zero_pivot_gives_zero_matrix():
    A = np.array([[2.0, 0.0, 1.0], [1.0, 0.0, 3.0], [0.0, 0.0, 5.0]])   # zero column
    ok, Ai = oracle.inv(A)
    assert not ok and not np.any(Ai)


def test_mtm_and_jtr():
    J = jacobian(2, 300, 40)
    r = np.random.default_rng(3).normal(size=300)
    JJ = oracle.mtm(J)
    assert np.array_equal(JJ, JJ.T)
    assert relerr(JJ, J.T @ J) < 1e-13
    assert relerr(oracle.jtr(J, r), J.T @ r) < 1e-13


def test_gn_steps_match_solve():
    rng = np.random.default_rng(4)
    J = jacobian(5, 400, 50)
    r = rng.normal(size=400)
    sig = rng.uniform(0.05, 4.0, 50)
    b = rng.normal(size=50)
    JJ, g = oracle.mtm(J), oracle.jtr(J, r)
    B = oracle.gn_steps(JJ, g, sig, DM_REF, b)
    for k, (d0, d1) in enumerate(DM_REF):
        K = JJ + d0 * np.diag(np.diag(JJ)) + d1 * np.diag(1.0 / sig**2)
        step = np.linalg.solve(K, g)
        assert np.max(np.abs(B[k] - b - step)) <= 1e-10 * np.max(np.abs(step)), k


def test_gn_steps_singular_is_no_step():
    """A parameter no datum sees (zero Jacobian column) with no damping at all:
    the reference's inverse is its zero matrix, so the step is 0."""
    J = jacobian(6, 100, 8)
    J[:, 3] = 0.0
    JJ, g = oracle.mtm(J), oracle.jtr(J, np.ones(100))
    b = np.arange(8.0)
    B = oracle.gn_steps(JJ, g, np.ones(8), np.array([[0.0, 0.0], [0.3, 1e4]]), b)
    assert np.array_equal(B[0], b)
    assert not np.array_equal(B[1], b)


def test_inv_diag():
    J = jacobian(7, 200, 30)
    JJ = oracle.mtm(J)
    for d in (1e-3, 1.0, 10.0):
        want = np.diag(np.linalg.inv(JJ + d * np.eye(30)))
        assert relerr(oracle.inv_diag(JJ, d), want) < 1e-10
    J[:, 5] = 0.0
    assert not np.any(oracle.inv_diag(oracle.mtm(J), 0.0))   # singular: zeros


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref (the reference's zmath.cpp) "
                    "is not built here")
def test_restatements_equal_the_reference_zmath():
    """oracle_mtm and oracle_inv against zmath.cpp's MTM and inv compiled from
    the reference sources: bit for bit (same operations, same order)."""
    ref = oracle.ref_lib()
    rng = np.random.default_rng(8)
    for rows, cols in ((50, 7), (333, 41)):
        J = jacobian(9 + cols, rows, cols)
        out = np.empty((cols, cols))
        ref.ref_mtm(rows, cols, oracle._ptr(J), oracle._ptr(out))
        assert np.array_equal(oracle.mtm(J), out)
    for A in (rng.normal(size=(23, 23)) + 3 * np.eye(23),
              np.array([[1e-14, 2.0, 0.0], [3.0, 1.0, 1.0], [0.0, 1.0, 4.0]]),
              oracle.mtm(jacobian(10, 80, 12)) + 0.3 * np.eye(12),
              np.array([[2.0, 0.0, 1.0], [1.0, 0.0, 3.0], [0.0, 0.0, 5.0]])):
        A = np.ascontiguousarray(A)
        out = np.empty_like(A)
        nz = ref.ref_inv(A.shape[0], oracle._ptr(A), oracle._ptr(out))
        ok, mine = oracle.inv(A)
        assert np.array_equal(mine, out)
        assert ok == bool(nz)
