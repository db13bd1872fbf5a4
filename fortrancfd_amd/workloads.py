"""Synthetic workloads of the benchmark and of the tests that run them at
their real sizes: the 3D rising bubble, the standalone variable-coefficient
Poisson problem, the batched 2D rising-bubble cases and the two-bubble
MPLS_RisingBubble problem.  Inputs are
deterministic (seeded numpy generators); the fluid starts at rest.
"""
from __future__ import annotations

import numpy as np

from .types import Case

BATCH_SIZES = ((64, 128), (128, 256), (256, 512))
SEED = 0x5EEDF00D


def rising_case(n: int, precision: int, tol: float) -> Case:
    """3D rising bubble on [0,1]^2 x [0,2] with n^3 cells (dz = 2 dx): density
    ratio 1000, viscosity ratio 100, CSF, gravity -z, no-slip walls,
    dt = 0.1 dx, reinitialisation every 5 steps, 4 Picard iterations."""
    return Case(dim=3, nx=n, ny=n, nz=n, xend=1.0, yend=1.0, zend=2.0, gvec=(0.0, 0.0, -1.0),
                cfl=0.1, cfl_r=0.1, ppe_tol=tol, ppe_precision=precision, redis_interval=5,
                picard_iters=4, ppe_max_cycles=100, ppe_rho_half=1, name="rising_bubble_3d")


def poisson_case(n: int, **kw) -> Case:
    """the standalone Poisson problem's grid: n^3 cells on the unit cube."""
    kw.setdefault("ppe_max_cycles", 200)
    kw.setdefault("name", "poisson")
    return Case(dim=3, nx=n, ny=n, nz=n, xend=1.0, yend=1.0, zend=1.0, **kw)


def sphere_density(case: Case) -> np.ndarray:
    """rho = 1000 outside and 1 inside a sphere of r = 0.3 at the centre,
    joined by the smoothed Heaviside of others.f90:14-25 over 1.5 cells; the
    full ghosted array."""
    X, Y, Z = case.coords()
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    del X, Y, Z
    x = (r - 0.3) / (1.5 * case.dx)
    del r
    H = np.where(x > 1, 1.0, np.where(x < -1, 0.0, 0.5 * (1 + x + np.sin(np.pi * x) / np.pi)))
    del x
    return np.ascontiguousarray(1.0 + 999.0 * H)


def seeded_rhs(case: Case, seed: int = SEED) -> np.ndarray:
    """seeded U[-1, 1] right-hand side whose interior mean is zero (ghosts
    are not read)."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1.0, 1.0, size=case.shape)
    fi = case.interior(f)
    fi -= fi.mean()
    return f


def batch_cases(ncases: int = 64, seed: int = SEED, precision: int = 1, tol: float = 1e-10):
    """Seeded batched rising-bubble cases on [0,1]x[0,2]: grid from
    64x128 / 128x256 / 256x512, bubble r ~ U[0.15, 0.3] at (0.5, 0.5),
    density ratio ~ U[10, 1000], Eotvos number ~ U[1, 100] with We = Eo
    (L = D, Fr = 1), viscosity ratio 100, Re = 1e4, reinit every 5 steps,
    PPE to tol (mixed: FP32 V-cycles in FP64 flexible CG).  The cases of
    one grid share dt (one context): dt = min(0.1 dx, half the explicit
    capillary limit sqrt(rho_eff We dx^3 / (2 pi)) of every case of that
    grid), with rho_eff = (1 - rho_12) / ln(1 / rho_12) the delta-weighted
    harmonic mean of the density across the smeared interface (the CSF
    force is divided by the local density).  We down to 1 is far below the
    CSF stability bound at dt = 0.1 dx; with the plain mean density as
    rho_eff a case with density ratio 747 and We = 11.7 grew parasitic
    currents to |u| ~ 10 and failed after ~160 steps.
    Returns [(case, bubble radius)]."""
    rng = np.random.default_rng(seed)
    draw = []
    for q in range(ncases):
        nx, ny = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        r = float(rng.uniform(0.15, 0.3))
        ratio = float(rng.uniform(10.0, 1000.0))
        eo = float(rng.uniform(1.0, 100.0))
        draw.append((nx, ny, r, 1.0 / ratio, eo))
    cfl = {}
    for nx, ny, r, r12, we in draw:
        dx = 1.0 / nx
        rho_eff = (1.0 - r12) / np.log(1.0 / r12)
        lim = 0.5 * np.sqrt(rho_eff * we * dx ** 3 / (2.0 * np.pi)) / dx
        cfl[nx] = min(cfl.get(nx, 0.1), lim)
    out = []
    for q, (nx, ny, r, r12, we) in enumerate(draw):
        c = Case(dim=2, nx=nx, ny=ny, nz=1, xend=1.0, yend=2.0, gvec=(0.0, -1.0, 0.0), rho_12=r12,
                 mu_12=0.01, re=1.0e4, we=we, fr=1.0, cfl=float(cfl[nx]), cfl_r=0.1, redis_interval=5,
                 ppe_tol=tol, ppe_precision=precision, name=f"batch{q}")
        out.append((c, r))
    return out


def m3_case(unit_grid: int, precision: int = 1, tol: float = 1e-10) -> Case:
    """The MPLS_RisingBubble problem (`MPLS_RisingBubble/20190612/SETUP.F90:14-85`):
    a [0,4] x [0,4] x [0,8] box with unit_grid cells per unit length (35 in
    the reference: 140 x 140 x 280), density / viscosity ratio 0.001 / 0.01,
    Re = 67.27, We = 16, Fr = 1, surface tension and gravity (-z) on,
    dt = 0.005 dx, rdt = 0.5 dx, interface half-width 1.5 dx, no-slip walls;
    the pressure to ||r||/||f|| <= tol."""
    n = int(unit_grid)
    return Case(dim=3, nx=4 * n, ny=4 * n, nz=8 * n, xend=4.0, yend=4.0, zend=8.0, cfl=0.005, cfl_r=0.5,
                ls_w=1.5, rho_12=0.001, mu_12=0.01, re=67.27, we=16.0, fr=1.0, btn_sf=1.0, btn_g=1.0,
                gvec=(0.0, 0.0, -1.0), ppe_tol=tol, ppe_precision=precision, ppe_max_cycles=100,
                name="mpls_rising_bubble")


def m3_bubbles(case: Case) -> dict:
    """the two bubbles of SETUP.F90:174-182: spheres of radius 0.5 at
    (2, 2, 1) and (2, 2, 2.5), phi > 0 inside; phi = max(phi1, phi2).  Full
    ghosted arrays (cell centres (I - 0.5) dx continued into the ghosts)."""
    X, Y, Z = case.coords()
    p1 = -np.sqrt((X - 2.0) ** 2 + (Y - 2.0) ** 2 + (Z - 1.0) ** 2) + 0.5
    p2 = -np.sqrt((X - 2.0) ** 2 + (Y - 2.0) ** 2 + (Z - 2.5) ** 2) + 0.5
    return {"phi1": np.ascontiguousarray(p1), "phi2": np.ascontiguousarray(p2),
            "phi": np.ascontiguousarray(np.maximum(p1, p2))}
