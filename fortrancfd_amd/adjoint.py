"""Differentiable level-set advection: the WENO5 SSP-RK3 steps of the
device path (fcfd_level_set_rk3_solver) behind a torch.autograd.Function
whose backward runs the hand-written discrete adjoint
(fcfd_ls_adv_adjoint_step) with checkpoint-and-recompute.

The prescribed velocity is the reference's 3D deformation field
(`Two phase flows/newOpenMP/3D+IBM/src/setup/velocity.f90:24-26`, cell
centred nvel, zero-gradient ghosts as its bc calls at :33-35), scaled by an
amplitude A and modulated by cos(pi t / T):
    u =  2 A sin^2(pi x) sin(2 pi y) sin(2 pi z) cos(pi t / T)
    v = -  A sin(2 pi x) sin^2(pi y) sin(2 pi z) cos(pi t / T)
    w = -  A sin(2 pi x) sin(2 pi y) sin^2(pi z) cos(pi t / T)
Step n advances from t_n = n dt with the velocity frozen at t_n (the
reference's level_set_rk3_solver uses nvel%old over the whole step).

Gradients: with respect to the initial level set (all interior cells) and
to A.  Since the velocity of step n is A ct_n uhat,
dJ/dA = sum_n ct_n <ubar_n, uhat>.  The forward keeps the level set of every
k-th step; the backward restores each segment from its checkpoint, re-runs
it storing its states, and walks the adjoint steps back through it.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .solver import Solver
from .types import Case, bc_copy


def deformation_field(case: Case) -> Tuple[np.ndarray, ...]:
    """Spatial part (A = 1, cos = 1) of the deformation velocity at the cell
    centres of the ghosted box, ghosts by zero-gradient copy."""
    if case.dim != 3:
        raise ValueError("the deformation field is the 3D one")
    X, Y, Z = case.coords()
    pi = math.pi
    u = 2.0 * np.sin(pi * X) ** 2 * np.sin(2 * pi * Y) * np.sin(2 * pi * Z)
    v = -np.sin(2 * pi * X) * np.sin(pi * Y) ** 2 * np.sin(2 * pi * Z)
    w = -np.sin(2 * pi * X) * np.sin(2 * pi * Y) * np.sin(pi * Z) ** 2
    return tuple(bc_copy(case, np.ascontiguousarray(a)) for a in (u, v, w))


class _Run:
    """Device state shared by the forward and backward of one problem."""

    def __init__(self, solver: Solver, nsteps: int, period: float, every: int):
        import torch
        self.s, self.case = solver, solver.case
        self.nsteps, self.period, self.every = int(nsteps), float(period), max(1, int(every))
        dev = torch.device("cuda")
        self.uhat = [torch.from_numpy(a.reshape(-1)).to(dev) for a in deformation_field(self.case)]
        self.vel = [torch.empty_like(a) for a in self.uhat]

    def ct(self, n: int) -> float:
        return math.cos(math.pi * n * self.case.dt / self.period)

    def set_velocity(self, A: float, n: int) -> None:
        import torch
        f = A * self.ct(n)
        for dst, src in zip(self.vel, self.uhat):
            torch.mul(src, f, out=dst)
        torch.cuda.synchronize()
        for name, t in zip(("nu", "nv", "nw"), self.vel):
            self.s.push_device(name, t.data_ptr())

    def step(self, A: float, n: int) -> None:
        self.set_velocity(A, n)
        self.s.level_set_rk3_solver()

    def pull_phi(self, out) -> None:
        self.s.pull_device("phi", out.data_ptr())
        self.s.sync()

    def push_phi(self, t) -> None:
        import torch
        torch.cuda.synchronize()
        self.s.push_device("phi", t.data_ptr())


def _function():
    import torch

    class LevelSetAdvection(torch.autograd.Function):
        """phi_T = advect(phi_0, A) over nsteps steps; see the module docs."""

        @staticmethod
        def forward(ctx, phi0, A, run: _Run):
            a = float(A.item())
            run.push_phi(phi0.detach().contiguous())
            ckpts: List = []
            for n in range(run.nsteps):
                if n % run.every == 0:
                    c = torch.empty_like(phi0)
                    run.pull_phi(c)
                    ckpts.append(c)
                run.step(a, n)
            out = torch.empty_like(phi0)
            run.pull_phi(out)
            ctx.run, ctx.ckpts, ctx.A = run, ckpts, a
            return out

        @staticmethod
        def backward(ctx, g):
            run, a = ctx.run, ctx.A
            lam = g.detach().clone().contiguous()
            bars = [torch.zeros_like(u) for u in run.uhat]
            gA = 0.0
            nseg = len(ctx.ckpts)
            for seg in range(nseg - 1, -1, -1):
                start = seg * run.every
                end = min(start + run.every, run.nsteps)
                # re-run the segment from its checkpoint, keeping its states
                run.push_phi(ctx.ckpts[seg])
                states = [ctx.ckpts[seg]]
                for n in range(start, end - 1):
                    run.step(a, n)
                    t = torch.empty_like(lam)
                    run.pull_phi(t)
                    states.append(t)
                for n in range(end - 1, start - 1, -1):
                    run.set_velocity(a, n)
                    for b in bars:
                        b.zero_()
                    torch.cuda.synchronize()
                    run.s.ls_adv_adjoint_step(states[n - start].data_ptr(), lam.data_ptr(),
                                              *(b.data_ptr() for b in bars))
                    run.s.sync()
                    gA += run.ct(n) * float(sum((b * u).sum() for b, u in zip(bars, run.uhat)))
            return lam, torch.tensor(gA, dtype=torch.float64, device=g.device), None

    return LevelSetAdvection


_LSA = None


def advect(phi0, A, solver: Solver, nsteps: int, period: float, every: int = 20):
    """Differentiable advection of phi0 (a CUDA float64 tensor of the ghosted
    field, flat or shaped) by the deformation field of amplitude A (a 0-d
    CUDA float64 tensor) over nsteps steps of solver.case.dt; checkpoints
    every `every` steps.  Returns phi_T with the shape of phi0."""
    global _LSA
    if _LSA is None:
        _LSA = _function()
    run = _Run(solver, nsteps, period, every)
    shape = phi0.shape
    out = _LSA.apply(phi0.reshape(-1), A, run)
    return out.reshape(shape)


def l2_loss(phi_T, target, case: Case):
    """J = ( sum over interior cells of (phi_T - target)^2 dV )^(1/2)."""
    g = case.ghc
    d = (phi_T - target).reshape(case.shape)[g:-g, g:-g, g:-g]
    return ((d * d).sum() * (case.dx * case.dy * case.dz)).sqrt()
