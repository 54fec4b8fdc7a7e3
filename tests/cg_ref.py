"""CPU checker for the batched solver (test infrastructure): conjugate gradients
on the regularised normal equations (G^T G + lam L^T L) x = G^T d, x0 = 0, in
float64, with G applied through caller-supplied functions, and L^T L for
forward differences inside a grid (no differences across the boundary)."""
import numpy as np


def ltl(v: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """(L^T L v) for v in raster order n = i + nx*(j + ny*k)."""
    a = np.asarray(v, np.float64).reshape(nz, ny, nx)
    out = np.zeros_like(a)
    for ax in range(3):
        d = np.diff(a, axis=ax)                      # v[i+1] - v[i] per edge
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        out[tuple(lo)] -= d
        out[tuple(hi)] += d
    return out.ravel()


def cg_normal(fwd, adj, reg, d, lam, iters):
    """Returns x and rr[k] = ||G^T d - A x_k||^2 for k = 0..iters, computed by
    the recurrences (rr[k+1] = r.r after the update of step k)."""
    b = adj(np.asarray(d, np.float64))
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = [float(r @ r)]
    for _ in range(iters):
        ap = adj(fwd(p)) + lam * reg(p)
        pap = float(p @ ap)
        alpha = rr[-1] / pap if (pap > 0 and rr[-1] > 0) else 0.0
        x += alpha * p
        r -= alpha * ap
        rn = float(r @ r)
        beta = rn / rr[-1] if rr[-1] > 0 else 0.0
        p = r + beta * p
        rr.append(rn)
    return x, np.array(rr)
