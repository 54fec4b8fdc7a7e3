"""One Levenberg-Marquardt iteration of ZTM's gravity fit on the device, the
caller side of the gravity forward model (LMA::fit, TFM.cpp:157-161, over
the forward model of TFM.cpp:83-99), built from the library's entry points:

    forward models     ztm_tfm_gravity_batched (producer + gravity, B variants)
    Jacobian columns   ztm_tfm_gravity_jvp (vm, ro, dr) and
                       ztm_tfm_gravity_jvp_grids (Moho basis coefficients)
    normal equations   ztm_gn_normal
    damped steps       ztm_gn_damped_steps (every damping condition at once)

The parameters are the reference's three scalars (vm, ro, dr) and the
coefficients of a caller-given basis of Moho grid perturbations (k units).
Where LMA::fit forms the columns by finite differences of whole forward
models (Jac, TFM.cpp:124-135), here they are the exact derivatives of the
producer's piecewise-linear density field; the candidates are compared by
their misfit as LMA::fit compares FM(S, B[n]).ssq."""
from __future__ import annotations

import numpy as np

from . import gfm


class GravityLM:
    """The gravity forward model of one ZTM geometry (an object with the
    fields of workloads.TfmInputs: bfr, di, dj, dk, fh, fv, rc, sea, moho,
    vel, surf) over the parameters b = (vm, ro, dr, c_1 .. c_K), the Moho
    grid being moho + sum_k c_k basis[k]."""

    def __init__(self, w, basis, device):
        import torch
        self.w = w
        self.dev = device
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(device)
        self.geo = (w.bfr, w.di, w.dj, w.dk, w.fh, w.fv, w.rc)
        self.sea, self.vel, self.moho = t64(w.sea), t64(w.vel), t64(w.moho)
        self.basis = t64(np.atleast_2d(basis))
        if self.basis.shape[1] != w.di * w.dj:
            raise ValueError("basis rows must be Moho grids of di * dj entries")
        self.s32 = torch.from_numpy(np.ascontiguousarray(w.surf, np.float32)).to(device)
        self.n = 3 + self.basis.shape[0]
        self._t64 = t64

    def forward(self, bs, stream=None):
        """Gravity [B, di*dj] (float64, exterior stations 0) of the parameter
        rows bs [B, n] (host array or device tensor)."""
        import torch
        bs = bs if isinstance(bs, torch.Tensor) else self._t64(np.atleast_2d(bs))
        if bs.dim() == 1:
            bs = bs[None, :]
        if bs.shape[1] != self.n:
            raise ValueError(f"parameter rows need {self.n} entries")
        moho = (self.moho + bs[:, 3:] @ self.basis).contiguous()
        g, _ = gfm.tfm_gravity_batched(*self.geo, self.sea, moho, self.vel,
                                       bs[:, :3].contiguous(), self.s32, stream=stream)
        return g

    def jacobian(self, b, stream=None):
        """dG/db at b: [di*dj, n] float64 (row-major, stations x parameters)."""
        import torch
        b = np.asarray(b, np.float64)
        moho = (self.moho + self._t64(b[3:]) @ self.basis).contiguous()
        dg3, _ = gfm.tfm_gravity_jvp(*self.geo, self.sea, moho, self.vel, b[0], b[1], b[2],
                                     self.s32, stream=stream)
        cols = [dg3]
        if self.n > 3:
            dgk, _ = gfm.tfm_gravity_jvp_grids(*self.geo, self.sea, moho, self.vel, b[0],
                                               b[1], b[2], self.s32,
                                               dmoho=self.basis.contiguous(), stream=stream)
            cols.append(dgk)
        return torch.cat(cols).T.contiguous()

    def iteration(self, b, d, sig, dm, stream=None):
        """One iteration from b against data d (device [di*dj]): returns
        (candidates B [nd, n], their misfits [nd], the index of the smallest,
        the misfit at b), all device tensors but the index."""
        import torch
        r = d - self.forward(b, stream)[0]
        J = self.jacobian(b, stream)
        JJ, jtr = gfm.gn_normal(J, r, stream=stream)
        B = gfm.gn_damped_steps(JJ, jtr, self._t64(sig), self._t64(dm), self._t64(b),
                                stream=stream)
        g = self.forward(B, stream)
        misfits = torch.sum((d[None, :] - g) ** 2, dim=1)
        return B, misfits, int(torch.argmin(misfits)), torch.sum(r * r)

    def fit(self, b, d, sig, dm, iters=4, g=0.999, stream=None):
        """LMA::fit's loop (TFM.cpp:157-166): up to `iters` iterations from
        b; each keeps the candidate of smallest misfit only if it is below
        the current misfit (the first such condition on ties, as the
        reference's in-order scan), logs the misfit, and the loop ends once
        misfit / previous misfit exceeds g. Returns (b, the misfit log Q,
        the index of the last accepted damping condition, or len(dm) - 1 if
        none was)."""
        b = np.asarray(b, np.float64).copy()
        if b.shape != (self.n,):
            raise ValueError(f"b needs {self.n} entries")
        Q, nb = [], len(dm) - 1
        for _ in range(iters):
            B, misfits, best, q0 = self.iteration(b, d, sig, dm, stream)
            if not Q:
                Q.append(float(q0))
            q, qt = Q[-1], float(misfits[best])
            if qt < q:
                q, b, nb = qt, B[best].cpu().numpy(), best
            Q.append(q)
            if Q[-2] == 0 or Q[-1] / Q[-2] > g:
                break
        return b, Q, nb
