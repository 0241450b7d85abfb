"""Gradient of the mean drag with respect to wall blowing / suction, by the discrete adjoint of
the batched 2D step (fvm_batch2d_adjoint_*), behind a torch.autograd.Function.

The control of a case is a vector of wall-normal velocities at equally spaced angular stations
around the cylinder; the wall cells J = 1..NJ-1 take the periodic linear interpolation of the
stations at their index position, along the outward normal (cos THETA, sin THETA) of the wall
cell (THETA(1, J) of the 2D reference's INIT).  The velocity enters the reference's wall hook
U_F, V_F of BCU / BCV (bcu.f90:10-20, bcv.f90:10-20).

The objective of a case is the mean of C_D (calccoef.f90:46-119) over the last navg of nsteps
steps.  Its gradient runs the adjoint step backwards over the trajectory with sqrt(N)
checkpointing: the forward run keeps the state at every k-th step (k = ceil(sqrt(N))) in device
slots; backwards, each segment is recomputed from its checkpoint with every state kept, then the
adjoint steps through it in reverse.  Device memory: about 2 sqrt(N) states.
"""
from __future__ import annotations

import math

import numpy as np

from .batch2d import Batch2D


class WallControl:
    """Maps the nctrl station velocities of one case to the wall velocities U_F, V_F (NJ+1)."""

    def __init__(self, nj: int, theta_wall: np.ndarray, nctrl: int = 128):
        self.nj, self.nctrl = int(nj), int(nctrl)
        n = nj - 1
        m = np.zeros((nj + 1, nctrl))
        for j in range(1, nj):
            s = (j - 1) * nctrl / n
            k0 = int(math.floor(s))
            f = s - k0
            m[j, k0 % nctrl] += 1.0 - f
            m[j, (k0 + 1) % nctrl] += f
        self.m = m
        th = np.asarray(theta_wall, dtype=np.float64).reshape(nj + 1)
        self.cos = np.where((np.arange(nj + 1) >= 1) & (np.arange(nj + 1) <= nj - 1), np.cos(th), 0.0)
        self.sin = np.where((np.arange(nj + 1) >= 1) & (np.arange(nj + 1) <= nj - 1), np.sin(th), 0.0)

    def velocities(self, ctrl):
        q = self.m @ np.asarray(ctrl, dtype=np.float64).reshape(self.nctrl)
        return q * self.cos, q * self.sin

    def pullback(self, gu, gv):
        """d/d(ctrl) from the gradients with respect to U_F, V_F."""
        return self.m.T @ (np.asarray(gu) * self.cos + np.asarray(gv) * self.sin)


class DragProblem:
    """Mean-drag objective of every case of a batch under wall control.

    batch: a Batch2D; x0: per case the initial (u, v, p, nlx, nly) planes; nsteps, navg: the
    horizon and the averaging window; tol / maxpass: the adjoint Poisson solve's scaled
    residual and pass cap."""

    def __init__(self, batch: Batch2D, x0, nsteps: int, navg: int, nctrl: int = 128,
                 tol: float = 1e-13, maxpass: int = 60):
        if not (1 <= navg <= nsteps):
            raise ValueError("need 1 <= navg <= nsteps")
        self.b, self.x0 = batch, [tuple(np.asarray(a) for a in x) for x in x0]
        self.nsteps, self.navg, self.tol, self.maxpass = int(nsteps), int(navg), tol, maxpass
        self.k = int(math.ceil(math.sqrt(self.nsteps)))
        self.nseg = (self.nsteps + self.k - 1) // self.k
        self.walls = [WallControl(c["nj"], batch.get(q, "theta")[1, :], nctrl)
                      for q, c in enumerate(batch.cases)]
        self._ctrl = None
        self.worst_ares = 0.0

    def _apply(self, ctrl):
        ctrl = np.asarray(ctrl, dtype=np.float64).reshape(self.b.n, -1)
        for q, w in enumerate(self.walls):
            self.b.set_wall_velocity(q, *w.velocities(ctrl[q]))
        return ctrl

    def _reset(self):
        for q, x in enumerate(self.x0):
            self.b.set_state(q, *x)

    def run(self, ctrl, keep=True):
        """The objective per case (mean C_D over the window); with keep, the segment
        checkpoints are stored for a gradient at the same control."""
        ctrl = self._apply(ctrl)
        self._reset()
        n = 0
        self.max_resip = 0.0
        while n < self.nsteps:  # whole segments between checkpoints, C_D from the history
            if keep and n % self.k == 0:
                self.b.save(n // self.k)
            m = min(self.k - n % self.k, self.nsteps - n)
            info = self.b.step(m)
            self.max_resip = max(self.max_resip, max(i["resip"] for i in info))
            n += m
        h = self.b.history(self.navg)  # (navg, ncase, 8); C_D first
        self._ctrl = ctrl.copy() if keep else self._ctrl
        return h[:, :, 0].sum(axis=0) / self.navg

    def gradient(self, ctrl, scale=None):
        """d(sum_c scale[c] J_c)/d(ctrl) as an (ncase, nctrl) array (scale defaults to ones)."""
        ctrl = np.asarray(ctrl, dtype=np.float64).reshape(self.b.n, -1)
        if self._ctrl is None or not np.array_equal(ctrl, self._ctrl):
            self.run(ctrl, keep=True)
        else:
            self._apply(ctrl)
        scale = np.ones(self.b.n) if scale is None else np.asarray(scale, dtype=np.float64)
        b = self.b
        b.adjoint_init()
        self.worst_ares = 0.0
        base = self.nseg  # slots after the checkpoints hold the states of one segment
        for seg in range(self.nseg - 1, -1, -1):
            s0, s1 = seg * self.k, min((seg + 1) * self.k, self.nsteps)
            b.load(seg)
            for m in range(s0, s1):  # recompute: keep the state before each step
                b.save(base + m - s0)
                if m < s1 - 1:
                    b.step(1)
            for n in range(s1, s0, -1):  # adjoint of step n (state n-1 -> n)
                b.load(base + n - 1 - s0)
                w = scale / self.navg if n > self.nsteps - self.navg else np.zeros(b.n)
                self.worst_ares = max(self.worst_ares, b.adjoint_step(w, self.tol, self.maxpass))
        g = np.zeros_like(ctrl)
        for q, wc in enumerate(self.walls):
            g[q] = wc.pullback(*b.adjoint_wall(q))
        return g


def mean_drag(problem: DragProblem, ctrl):
    """The objective per case as a differentiable torch function of ctrl (ncase x nctrl,
    float64)."""
    import torch

    class MeanDrag(torch.autograd.Function):
        @staticmethod
        def forward(ctx, c):
            ctx.save_for_backward(c)
            j = problem.run(c.detach().cpu().numpy(), keep=True)
            return torch.as_tensor(j, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gj):
            (c,) = ctx.saved_tensors
            g = problem.gradient(c.detach().cpu().numpy(), scale=gj.detach().cpu().numpy())
            return torch.as_tensor(g, dtype=c.dtype).reshape(c.shape)

    return MeanDrag.apply(ctrl)
