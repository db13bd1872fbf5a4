"""Independent FP64 reference gradient for the level-set advection adjoint
(test infrastructure): the WENO5 flux-split SSP-RK3 step restated with
torch tensor operations on the CPU, differentiated by torch.autograd.

It follows the same reference formulas as the device forward
(`Two phase flows/newOpenMP/3D+IBM/src/interface_solver/
level_set_rk3_solver.f90`: flux split :98-160, WENO5-JS :162-219 with
EPS = 1e-13 and 12x the Jiang-Shu smoothness indicators, increment-form
RK3 :1-65), with every stencil position clamped to the interior as the
reference's zero-gradient ghosts imply.  Arrays are interior-only,
indexed [k, j, i].
"""
import math

import torch

EPS = 1.0e-13


def _weights(b1, b2, b3):
    q1, q2, q3 = (EPS + b1) ** 2, (EPS + b2) ** 2, (EPS + b3) ** 2
    a1, a2, a3 = 1.0 / q1, 6.0 / q2, 3.0 / q3
    s = a1 + a2 + a3
    return a1 / s, a2 / s, a3 / s


def _left(g0, g1, g2, g3, g4):
    b1 = 13.0 * (g0 - 2 * g1 + g2) ** 2 + 3.0 * (g0 - 4 * g1 + 3 * g2) ** 2
    b2 = 13.0 * (g1 - 2 * g2 + g3) ** 2 + 3.0 * (g1 - g3) ** 2
    b3 = 13.0 * (g2 - 2 * g3 + g4) ** 2 + 3.0 * (3 * g2 - 4 * g3 + g4) ** 2
    w1, w2, w3 = _weights(b1, b2, b3)
    return (w1 * (2 * g0 - 7 * g1 + 11 * g2) + w2 * (-g1 + 5 * g2 + 2 * g3) + w3 * (2 * g2 + 5 * g3 - g4)) / 6.0


def _right(f0, f1, f2, f3, f4):
    b3 = 13.0 * (f0 - 2 * f1 + f2) ** 2 + 3.0 * (f0 - 4 * f1 + 3 * f2) ** 2
    b2 = 13.0 * (f1 - 2 * f2 + f3) ** 2 + 3.0 * (f1 - f3) ** 2
    b1 = 13.0 * (f2 - 2 * f3 + f4) ** 2 + 3.0 * (3 * f2 - 4 * f3 + f4) ** 2
    w1, w2, w3 = _weights(b1, b2, b3)
    return (w3 * (-f0 + 5 * f1 + 2 * f2) + w2 * (2 * f1 + 5 * f2 - f3) + w1 * (11 * f2 - 7 * f3 + 2 * f4)) / 6.0


def _take(a, axis, idx):
    return torch.index_select(a, axis, idx)


def L(phi, vel, h):
    """-sum_d (F_{i+1/2} - F_{i-1/2}) / h_d on an interior [k, j, i] array;
    vel = (u, v, w) interior arrays, h = (dx, dy, dz)."""
    out = torch.zeros_like(phi)
    for d, axis in enumerate((2, 1, 0)):
        n = phi.shape[axis]
        u = vel[d]
        s1 = 0.5 * (u + u.abs()) * phi
        s2 = 0.5 * (u - u.abs()) * phi
        f = torch.arange(-1, n)                       # faces f + 1/2, f = -1 .. n-1
        cl = lambda o: torch.clamp(f + o, 0, n - 1)
        F = _left(*(_take(s1, axis, cl(o)) for o in (-2, -1, 0, 1, 2))) + \
            _right(*(_take(s2, axis, cl(o)) for o in (-1, 0, 1, 2, 3)))
        Fr = _take(F, axis, torch.arange(1, n + 1))   # face i + 1/2
        Fl = _take(F, axis, torch.arange(0, n))       # face i - 1/2
        out = out - (Fr - Fl) / h[d]
    return out


def rk3_step(phi, vel, h, dt):
    l0 = L(phi, vel, h)
    p1 = phi + dt * l0
    l1 = L(p1, vel, h)
    p2 = p1 + dt * (-3.0 * l0 + l1) / 4.0
    l2 = L(p2, vel, h)
    return p2 + dt * (-l0 - l1 + 8.0 * l2) / 12.0


def deformation(X, Y, Z):
    pi = math.pi
    return (2.0 * torch.sin(pi * X) ** 2 * torch.sin(2 * pi * Y) * torch.sin(2 * pi * Z),
            -torch.sin(2 * pi * X) * torch.sin(pi * Y) ** 2 * torch.sin(2 * pi * Z),
            -torch.sin(2 * pi * X) * torch.sin(2 * pi * Y) * torch.sin(pi * Z) ** 2)


def advect(phi0, A, uhat, h, dt, nsteps, period):
    phi = phi0
    for n in range(nsteps):
        ct = math.cos(math.pi * n * dt / period)
        phi = rk3_step(phi, tuple(A * ct * u for u in uhat), h, dt)
    return phi
