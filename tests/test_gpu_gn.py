"""The dense LA of the Gauss-Newton step on the GPU (ztm_gn_*) against the
oracle's restatements of zmath.cpp's MTM and inv and of TFM.cpp's damped
steps and inverse diagonal (FP64). The normal equations differ from the
oracle only in summation order (1e-13 of the largest entry); the steps and
the inverse diagonal are Cholesky solves against the reference's Gauss-Jordan
inverse, both FP64, held to 1e-9 of the largest step on these systems."""
import numpy as np
import pytest

from ztm_amd import gfm

pytestmark = pytest.mark.gpu

DM_REF = np.array([[a, b] for a in (0, .3, 1, 3, 10, 100) for b in (1e4, 1e5)])


def jacobian(seed, rows, cols):
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(rows, cols)) * np.exp(rng.uniform(-3, 3, cols))
    J[rng.uniform(size=(rows, cols)) < 0.3] = 0.0
    return J


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


@pytest.mark.parametrize("rows,n", [(1, 1), (17, 5), (300, 127), (300, 130), (1000, 257),
                                    (5000, 64)])
def test_normal_equations_vs_oracle(dev, orc, rows, n):
    print(f"[gn normal] rows={rows} n={n}", flush=True)
    J = jacobian(rows + n, rows, n)
    r = np.random.default_rng(rows).normal(size=rows)
    JJ, jtr = gfm.gn_normal(_t(J, dev), _t(r, dev))
    JJ, jtr = JJ.cpu().numpy(), jtr.cpu().numpy()
    assert relerr(JJ, orc.mtm(J)) < 1e-13
    assert relerr(jtr, orc.jtr(J, r)) < 1e-13
    JJ2, none = gfm.gn_normal(_t(J, dev))
    assert none is None and np.array_equal(JJ2.cpu().numpy(), JJ)


def test_normal_equations_no_rows(dev):
    import torch
    J = torch.zeros(0, 9, dtype=torch.float64, device=dev)
    JJ, jtr = gfm.gn_normal(J, torch.zeros(0, dtype=torch.float64, device=dev))
    assert not torch.any(JJ).item() and not torch.any(jtr).item()


@pytest.mark.parametrize("n", [1, 31, 32, 33, 64, 96, 100, 200, 257])
def test_damped_steps_vs_oracle(dev, orc, n):
    print(f"[gn steps] n={n}", flush=True)
    rng = np.random.default_rng(n)
    rows = 3 * n + 20
    J = jacobian(n, rows, n)
    r = rng.normal(size=rows)
    sig = rng.uniform(0.05, 4.0, n)
    b = rng.normal(size=n)
    JJ, jtr = orc.mtm(J), orc.jtr(J, r)
    want = orc.gn_steps(JJ, jtr, sig, DM_REF, b)
    B = gfm.gn_damped_steps(_t(JJ, dev), _t(jtr, dev), _t(sig, dev), _t(DM_REF, dev),
                            _t(b, dev)).cpu().numpy()
    for k in range(len(DM_REF)):
        step = want[k] - b
        assert np.max(np.abs(B[k] - want[k])) <= 1e-9 * np.max(np.abs(step)), k


def test_damped_steps_singular_system_gives_no_step(dev, orc):
    """A parameter no datum sees, undamped: the reference's zero inverse, no step."""
    J = jacobian(3, 120, 40)
    J[:, 17] = 0.0
    JJ, jtr = orc.mtm(J), orc.jtr(J, np.ones(120))
    b = np.arange(40.0)
    dm = np.array([[0.0, 0.0], [0.3, 1e4]])
    want = orc.gn_steps(JJ, jtr, np.ones(40), dm, b)
    B = gfm.gn_damped_steps(_t(JJ, dev), _t(jtr, dev), _t(np.ones(40), dev), _t(dm, dev),
                            _t(b, dev)).cpu().numpy()
    assert np.array_equal(B[0], b) and np.array_equal(want[0], b)
    assert np.max(np.abs(B[1] - want[1])) <= 1e-9 * np.max(np.abs(want[1] - b))


@pytest.mark.parametrize("n", [5, 64, 150])
def test_inverse_diag_vs_oracle(dev, orc, n):
    J = jacobian(n + 7, 4 * n, n)
    JJ = orc.mtm(J)
    for d in (1e-3, 1.0, 10.0):
        got = gfm.gn_inverse_diag(_t(JJ, dev), d).cpu().numpy()
        assert relerr(got, orc.inv_diag(JJ, d)) < 1e-9, d
    J[:, n // 2] = 0.0
    assert not np.any(gfm.gn_inverse_diag(_t(orc.mtm(J), dev), 0.0).cpu().numpy())


def test_gn_rejects(dev):
    import torch
    with pytest.raises(ValueError):
        gfm.gn_normal(torch.zeros(4, 3, dtype=torch.float32, device=dev))
    big = torch.zeros(7681, 7681, dtype=torch.float64, device=dev)
    with pytest.raises(gfm.ZtmError):
        gfm.gn_inverse_diag(big, 1.0)
    del big


def test_gn_full_size_properties(dev):
    """32,768 data x 1,024 parameters, the reference's 12 damping conditions:
    J^T J symmetric and J^T J x = J^T (J x); every step solves its system,
    |K_k x_k - J^T r| <= 1e-10 |J^T r|."""
    import torch
    rows, n = 32768, 1024
    J = _t(jacobian(11, rows, n), dev)
    r = _t(np.random.default_rng(12).normal(size=rows), dev)
    JJ, jtr = gfm.gn_normal(J, r)
    assert torch.equal(JJ, JJ.T)
    x = _t(np.random.default_rng(13).normal(size=n), dev)
    lhs, rhs = JJ @ x, J.T @ (J @ x)
    assert (torch.max(torch.abs(lhs - rhs)) / torch.max(torch.abs(rhs))).item() < 1e-12
    assert (torch.max(torch.abs(jtr - J.T @ r)) / torch.max(torch.abs(jtr))).item() < 1e-12
    sig = _t(np.random.default_rng(14).uniform(0.05, 4.0, n), dev)
    b = torch.zeros(n, dtype=torch.float64, device=dev)
    B = gfm.gn_damped_steps(JJ, jtr, sig, _t(DM_REF, dev), b)
    dJ = torch.diagonal(JJ)
    for k, (d0, d1) in enumerate(DM_REF):
        K = JJ + torch.diag(d0 * dJ + d1 / sig**2)
        res = torch.max(torch.abs(K @ B[k] - jtr)) / torch.max(torch.abs(jtr))
        assert res.item() < 1e-10, k


def test_damped_steps_many_conditions(dev, orc):
    """40 seeded damping conditions (d0 in [0, 100], d1 in [1e3, 1e5]) at
    n = 97 (three full 32-column blocks and a 1-column one), 300 data: every
    step against the oracle's Gauss-Jordan step to 1e-9 of its size."""
    rng = np.random.default_rng(97)
    n, rows, nd = 97, 300, 40
    J = jacobian(98, rows, n)
    r = rng.normal(size=rows)
    sig = rng.uniform(0.05, 4.0, n)
    b = rng.normal(size=n)
    d1 = np.exp(rng.uniform(np.log(1e3), np.log(1e5), nd))
    dm = np.stack([rng.uniform(0.0, 100.0, nd), d1], axis=1)
    JJ, jtr = orc.mtm(J), orc.jtr(J, r)
    want = orc.gn_steps(JJ, jtr, sig, dm, b)
    B = gfm.gn_damped_steps(_t(JJ, dev), _t(jtr, dev), _t(sig, dev), _t(dm, dev),
                            _t(b, dev)).cpu().numpy()
    for k in range(nd):
        assert np.max(np.abs(B[k] - want[k])) <= 1e-9 * np.max(np.abs(want[k] - b)), k


def test_past_1024_parameters_vs_oracle(dev, orc):
    """n = 1,100 (34 full 32-column blocks and a 12-column one): the
    substitutions' columns past the first 1,024 (read from K, not prefetched)
    and the update tiles of a ragged last block, against the oracle's
    Gauss-Jordan step and inverse diagonal."""
    n, rows = 1100, 2400
    print(f"[gn steps] n={n} rows={rows}", flush=True)
    rng = np.random.default_rng(1100)
    J = jacobian(1101, rows, n)
    r = rng.normal(size=rows)
    sig = rng.uniform(0.05, 4.0, n)
    b = rng.normal(size=n)
    dm = np.array([[0.3, 1e4]])
    JJ, jtr = orc.mtm(J), orc.jtr(J, r)
    want = orc.gn_steps(JJ, jtr, sig, dm, b)
    B = gfm.gn_damped_steps(_t(JJ, dev), _t(jtr, dev), _t(sig, dev), _t(dm, dev),
                            _t(b, dev)).cpu().numpy()
    assert np.max(np.abs(B[0] - want[0])) <= 1e-9 * np.max(np.abs(want[0] - b))
    got = gfm.gn_inverse_diag(_t(JJ, dev), 1.0).cpu().numpy()
    assert relerr(got, orc.inv_diag(JJ, 1.0)) < 1e-9


def test_random_shapes_vs_oracle(dev, orc):
    """A seeded sweep of 8 shapes (n in [1, 300], data in [n, 3n], 1 to 20
    damping conditions): the normal equations, every step and the inverse
    diagonal against the oracle."""
    rng = np.random.default_rng(2024)
    for case in range(8):
        n = int(rng.integers(1, 301))
        rows = int(rng.integers(n, 3 * n + 1))
        nd = int(rng.integers(1, 21))
        print(f"[gn sweep] case {case}: n={n} rows={rows} nd={nd}", flush=True)
        J = jacobian(case, rows, n)
        r = rng.normal(size=rows)
        sig = rng.uniform(0.05, 4.0, n)
        b = rng.normal(size=n)
        dm = np.stack([rng.uniform(0.0, 100.0, nd),
                       np.exp(rng.uniform(np.log(1e3), np.log(1e5), nd))], axis=1)
        JJd, jtrd = gfm.gn_normal(_t(J, dev), _t(r, dev))
        JJ, jtr = orc.mtm(J), orc.jtr(J, r)
        assert relerr(JJd.cpu().numpy(), JJ) < 1e-13, case
        assert relerr(jtrd.cpu().numpy(), jtr) < 1e-13, case
        want = orc.gn_steps(JJ, jtr, sig, dm, b)
        B = gfm.gn_damped_steps(_t(JJ, dev), _t(jtr, dev), _t(sig, dev), _t(dm, dev),
                                _t(b, dev)).cpu().numpy()
        for k in range(nd):
            assert np.max(np.abs(B[k] - want[k])) <= 1e-9 * np.max(np.abs(want[k] - b)), (case, k)
        d = float(rng.uniform(0.01, 10.0))
        got = gfm.gn_inverse_diag(_t(JJ, dev), d).cpu().numpy()
        assert relerr(got, orc.inv_diag(JJ, d)) < 1e-9, case
