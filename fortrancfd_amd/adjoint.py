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
to A.  The fields NU/NV/NW hold A uhat for the whole run and step n uses
them with the context's velocity scale ct_n (fcfd_set_velocity_scale), so
the adjoint steps accumulate dJ/d(A uhat) = sum_n ct_n dJ/du_n and
dJ/dA = <dJ/d(A uhat), uhat>.  The forward keeps the level set of every
k-th step; the backward restores each segment from its checkpoint, re-runs
it storing its states, and walks the adjoint steps back through it.

`advect` differentiates in the precision of phi0: float64 runs the FP64
steps and their adjoint; float32 the FP32 steps
(fcfd_level_set_rk3_solver_f32) and their FP32 adjoint
(fcfd_ls_adv_adjoint_step_f32), with the velocity and checkpoints in FP32.
`advect_f32` runs the FP32 forward alone (no gradient).
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


def deformation_field_device(case: Case, device, dtype=None):
    """deformation_field evaluated on the device with torch (flat ghosted
    arrays): every ghost cell takes the formula at its clamped interior
    index, which is the zero-gradient ghost copy."""
    import torch
    if case.dim != 3:
        raise ValueError("the deformation field is the 3D one")
    dtype = dtype or torch.float64
    g = case.ghc

    def centres(n, start, h):
        i = torch.clamp(torch.arange(-g, n + g, device=device, dtype=torch.float64), 0, n - 1)
        return start + (i + 0.5) * h

    x = centres(case.nx, case.xstart, case.dx).reshape(1, 1, -1)
    y = centres(case.ny, case.ystart, case.dy).reshape(1, -1, 1)
    z = centres(case.nz, case.zstart, case.dz).reshape(-1, 1, 1)
    pi = math.pi
    sx, sy, sz = torch.sin(pi * x), torch.sin(pi * y), torch.sin(pi * z)
    s2x, s2y, s2z = torch.sin(2 * pi * x), torch.sin(2 * pi * y), torch.sin(2 * pi * z)
    u = 2.0 * (sx * sx) * s2y * s2z
    v = -s2x * (sy * sy) * s2z
    w = -s2x * s2y * (sz * sz)
    return tuple(a.to(dtype).contiguous().reshape(-1) for a in (u, v, w))


class _Run:
    """Device state shared by the forward and backward of one problem."""

    def __init__(self, solver: Solver, nsteps: int, period: float, every: int):
        import torch
        self.s, self.case = solver, solver.case
        self.nsteps, self.period, self.every = int(nsteps), float(period), max(1, int(every))
        dev = torch.device("cuda")
        self.uhat = list(deformation_field_device(self.case, dev))
        self.vel = [torch.empty_like(a) for a in self.uhat]

    def ct(self, n: int) -> float:
        return math.cos(math.pi * n * self.case.dt / self.period)

    def set_amplitude(self, A: float) -> None:
        """NU/NV/NW = A uhat (the per-step factor goes in the velocity scale)."""
        import torch
        for dst, src in zip(self.vel, self.uhat):
            torch.mul(src, A, out=dst)
        torch.cuda.synchronize()
        for name, t in zip(("nu", "nv", "nw"), self.vel):
            self.s.push_device(name, t.data_ptr())

    def step(self, n: int) -> None:
        self.s.set_velocity_scale(self.ct(n))
        self.s.level_set_rk3_solver()

    def pull_phi(self, out) -> None:
        """Stream-ordered copy of the solver's phi into `out` (no wait)."""
        self.s.pull_device("phi", out.data_ptr())

    def push_phi(self, t) -> None:
        import torch
        torch.cuda.synchronize()
        self.s.push_device("phi", t.data_ptr())

    def finish(self) -> None:
        """Wait for the solver's stream and restore the unit velocity scale."""
        self.s.sync()
        self.s.set_velocity_scale(1.0)


def _fold_ghosts(a, g: int):
    """Adjoint of the zero-gradient ghost copy (bc): every ghost of the
    ghosted array `a` (shaped (k, j, i) or (j, i)) holds the value of the
    interior cell its indices clamp to, so its cotangent is added to that
    cell; the ghosts are then zero.  Axis by axis, so corner and edge ghosts
    reach their cell through the faces they are copied from."""
    a = a.clone()
    for ax in range(a.dim()):
        n = a.shape[ax]
        a.narrow(ax, g, 1).add_(a.narrow(ax, 0, g).sum(dim=ax, keepdim=True))
        a.narrow(ax, n - g - 1, 1).add_(a.narrow(ax, n - g, g).sum(dim=ax, keepdim=True))
        a.narrow(ax, 0, g).zero_()
        a.narrow(ax, n - g, g).zero_()
    return a


def _zero_ghosts(a, g: int):
    for ax in range(a.dim()):
        n = a.shape[ax]
        a.narrow(ax, 0, g).zero_()
        a.narrow(ax, n - g, g).zero_()
    return a


def _function():
    import torch

    class LevelSetAdvection(torch.autograd.Function):
        """phi_T = advect(phi_0, A) over nsteps steps; see the module docs."""

        @staticmethod
        def forward(ctx, phi0, A, run: _Run):
            a = float(A.item())
            run.set_amplitude(a)
            run.push_phi(phi0.detach().contiguous())
            ckpts: List = []
            for n in range(run.nsteps):
                if n % run.every == 0:
                    c = torch.empty_like(phi0)
                    run.pull_phi(c)
                    ckpts.append(c)
                run.step(n)
            out = torch.empty_like(phi0)
            run.pull_phi(out)
            run.finish()
            ctx.run, ctx.ckpts, ctx.A = run, ckpts, a
            return out

        @staticmethod
        def backward(ctx, g):
            run, a = ctx.run, ctx.A
            run.set_amplitude(a)
            case = run.s.case
            # phi_T's ghosts are zero-gradient copies made by the last stage:
            # their cotangents belong to the interior cells they copy
            lam = _fold_ghosts(g.detach().reshape(case.shape), case.ghc).reshape(-1).contiguous()
            bars = [torch.zeros_like(u) for u in run.uhat]   # dJ/d(A uhat), summed over the steps
            torch.cuda.synchronize()
            nseg = len(ctx.ckpts)
            for seg in range(nseg - 1, -1, -1):
                start = seg * run.every
                end = min(start + run.every, run.nsteps)
                # re-run the segment from its checkpoint, keeping its states
                run.push_phi(ctx.ckpts[seg])
                states = [ctx.ckpts[seg]]
                for n in range(start, end - 1):
                    run.step(n)
                    t = torch.empty_like(lam)
                    run.pull_phi(t)
                    states.append(t)
                for n in range(end - 1, start - 1, -1):
                    run.s.set_velocity_scale(run.ct(n))
                    run.s.ls_adv_adjoint_step(states[n - start].data_ptr(), lam.data_ptr(),
                                              *(b.data_ptr() for b in bars))
                run.s.sync()   # the segment's states are released after this
            run.finish()
            gA = sum((b * u).sum() for b, u in zip(bars, run.uhat))
            # the advection never reads phi0's ghosts
            lam = _zero_ghosts(lam.reshape(case.shape), case.ghc).reshape(-1)
            return lam, gA.reshape(()).to(torch.float64), None

    return LevelSetAdvection


class _Run32:
    """FP32 state of one problem: the float32 velocity A uhat lives in
    caller tensors (the FP32 step takes pointers), the level set in a
    tensor of the run."""

    def __init__(self, solver: Solver, nsteps: int, period: float, every: int):
        import torch
        self.s, self.case = solver, solver.case
        self.nsteps, self.period, self.every = int(nsteps), float(period), max(1, int(every))
        self.uhat = [u.to(torch.float32) for u in deformation_field_device(self.case, torch.device("cuda"))]
        self.vel = [torch.empty_like(a) for a in self.uhat]

    def ct(self, n: int) -> float:
        return math.cos(math.pi * n * self.case.dt / self.period)

    def set_amplitude(self, A: float) -> None:
        import torch
        for dst, src in zip(self.vel, self.uhat):
            torch.mul(src, A, out=dst)
        torch.cuda.synchronize()

    def ptrs(self):
        return [v.data_ptr() for v in self.vel]

    def step(self, n: int, phi) -> None:
        self.s.set_velocity_scale(self.ct(n))
        self.s.level_set_rk3_solver_f32(phi.data_ptr(), *self.ptrs())

    def finish(self) -> None:
        self.s.sync()
        self.s.set_velocity_scale(1.0)


def _function32():
    import torch

    class LevelSetAdvection32(torch.autograd.Function):
        """phi_T = advect(phi_0, A) in FP32; see the module docs."""

        @staticmethod
        def forward(ctx, phi0, A, run: _Run32):
            a = float(A.item())
            run.set_amplitude(a)
            phi = phi0.detach().contiguous().clone()
            ckpts: List = []
            torch.cuda.synchronize()
            for n in range(run.nsteps):
                if n % run.every == 0:
                    run.s.sync()
                    ckpts.append(phi.clone())
                    torch.cuda.synchronize()
                run.step(n, phi)
            run.finish()
            ctx.run, ctx.ckpts, ctx.A, ctx.A_dtype = run, ckpts, a, A.dtype
            return phi

        @staticmethod
        def backward(ctx, g):
            run, a = ctx.run, ctx.A
            run.set_amplitude(a)
            case = run.s.case
            lam = _fold_ghosts(g.detach().reshape(case.shape), case.ghc).reshape(-1).contiguous()
            bars = [torch.zeros_like(u) for u in run.uhat]
            torch.cuda.synchronize()
            for seg in range(len(ctx.ckpts) - 1, -1, -1):
                start = seg * run.every
                end = min(start + run.every, run.nsteps)
                states = [ctx.ckpts[seg]]
                cur = ctx.ckpts[seg].clone()
                torch.cuda.synchronize()
                for n in range(start, end - 1):
                    run.step(n, cur)
                    run.s.sync()
                    states.append(cur.clone())
                    torch.cuda.synchronize()
                for n in range(end - 1, start - 1, -1):
                    run.s.set_velocity_scale(run.ct(n))
                    run.s.ls_adv_adjoint_step_f32(states[n - start].data_ptr(), lam.data_ptr(),
                                                  *(b.data_ptr() for b in bars), *run.ptrs())
                run.s.sync()
            run.finish()
            gA = sum((b.double() * u.double()).sum() for b, u in zip(bars, run.uhat))
            lam = _zero_ghosts(lam.reshape(case.shape), case.ghc).reshape(-1)
            return lam, gA.reshape(()).to(ctx.A_dtype), None

    return LevelSetAdvection32


_LSA = None
_LSA32 = None


def advect(phi0, A, solver: Solver, nsteps: int, period: float, every: int = 20):
    """Differentiable advection of phi0 (a CUDA float64 or float32 tensor of
    the ghosted field, flat or shaped) by the deformation field of amplitude
    A (a 0-d CUDA tensor) over nsteps steps of solver.case.dt, in phi0's
    precision; checkpoints every `every` steps.  Returns phi_T with the shape
    and dtype of phi0."""
    import torch
    global _LSA, _LSA32
    shape = phi0.shape
    if phi0.dtype == torch.float32:
        if _LSA32 is None:
            _LSA32 = _function32()
        run = _Run32(solver, nsteps, period, every)
        out = _LSA32.apply(phi0.reshape(-1), A, run)
        return out.reshape(shape)
    if _LSA is None:
        _LSA = _function()
    run = _Run(solver, nsteps, period, every)
    out = _LSA.apply(phi0.reshape(-1), A, run)
    return out.reshape(shape)


def advect_f32(phi0, A: float, solver: Solver, nsteps: int, period: float):
    """The forward of `advect` in FP32: phi0 (CUDA tensor of the ghosted
    field) is rounded to float32, advanced nsteps WENO5 SSP-RK3 steps by
    fcfd_level_set_rk3_solver_f32 with the float32 velocity A uhat times
    cos(pi n dt / T), and returned as float32.  No gradient."""
    import torch
    case = solver.case
    dev = torch.device("cuda")
    phi = phi0.detach().reshape(-1).to(torch.float32).contiguous().clone()
    vel = [(a * float(A)).to(torch.float32) for a in deformation_field_device(case, dev)]
    ptrs = [v.data_ptr() for v in vel]
    torch.cuda.synchronize()
    for n in range(int(nsteps)):
        solver.set_velocity_scale(math.cos(math.pi * n * case.dt / float(period)))
        solver.level_set_rk3_solver_f32(phi.data_ptr(), *ptrs)
    solver.sync()
    solver.set_velocity_scale(1.0)
    return phi.reshape(phi0.shape)


def l2_loss(phi_T, target, case: Case):
    """J = ( sum over interior cells of (phi_T - target)^2 dV )^(1/2)."""
    g = case.ghc
    d = (phi_T - target).reshape(case.shape)[g:-g, g:-g, g:-g]
    return ((d * d).sum() * (case.dx * case.dy * case.dz)).sqrt()
