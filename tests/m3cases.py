"""Seeded states of the MPLS_RisingBubble (two-bubble) problem shared by the
M3 oracle-vs-reference tests and the M3 GPU parity tests.

The reference's own initial condition (SETUP.F90:174-200) is two spheres of
radius 0.5 centred on the vertical axis of a 4 x 4 x 8 box; here the same
construction is scaled to small boxes: PHI1, PHI2 are the signed distances
(positive inside, the bubble convention of AMURHO.F90:18-22), PHI their max.
Grids keep NODE_Y = NODE_Z, where the reference's z-direction CCD loops
(FIND_CURVATURE.F90:128,157, which run the y index to NODE_Z) are right.
"""
from __future__ import annotations

import numpy as np

from oracle.m3 import M3Case


def m3_case(n: int = 12, **kw) -> M3Case:
    """nx = n, ny = nz = 4n/3 cells on [0, 1.5] x [0, 2] x [0, 2] (dx = 1.5/n)."""
    m = (4 * n) // 3
    return M3Case(nx=n, ny=m, nz=m, dx=1.5 / n, **kw)


def bubbles(case: M3Case, r: float = 0.35, zc=(0.6, 1.4)) -> dict:
    X, Y, Z = case.coords()
    xc, yc = 0.5 * case.nx * case.dx, 0.5 * case.ny * case.dx
    p1 = -np.sqrt((X - xc) ** 2 + (Y - yc) ** 2 + (Z - zc[0]) ** 2) + r
    p2 = -np.sqrt((X - xc) ** 2 + (Y - yc) ** 2 + (Z - zc[1]) ** 2) + r
    return {"phi1": np.ascontiguousarray(p1), "phi2": np.ascontiguousarray(p2),
            "phi": np.ascontiguousarray(np.maximum(p1, p2))}


def smooth_velocity(case: M3Case, seed: int, amp: float = 0.3) -> dict:
    """u, v, w: a few seeded low modes, zero on the wall faces of their own
    direction (so the no-slip BC rules hold approximately)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = case.coords()
    L = (case.nx * case.dx, case.ny * case.dx, case.nz * case.dx)
    out = {}
    for name, shift in (("u", (0.5, 0, 0)), ("v", (0, 0.5, 0)), ("w", (0, 0, 0.5))):
        Xs, Ys, Zs = X + shift[0] * case.dx, Y + shift[1] * case.dx, Z + shift[2] * case.dx
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(1, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += rng.uniform(-1, 1) * np.sin(k[0] * np.pi * Xs / L[0] + ph[0]) * \
                np.sin(k[1] * np.pi * Ys / L[1] + ph[1]) * np.sin(k[2] * np.pi * Zs / L[2] + ph[2])
        out[name] = np.ascontiguousarray(amp * f)
    return out


def interior(case: M3Case, a):
    return case.interior(a)


def max_rel(a, b, case: M3Case = None) -> float:
    if case is not None:
        a, b = case.interior(a), case.interior(b)
    den = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) / den


# single operations pinned by the golden fixture: harness op -> fields kept
GOLDEN_OPS = {
    "amurho": ("rho", "amu", "heavy", "delta"),
    "ls_ccd": ("normal_x", "normal_z", "curv"),
    "ls_source": ("s0",),
    "ls_rk3": ("phi",),
    "redistance": ("phi",),
    "mpls1": ("phi",),
    "solve_ls": ("phi", "phi1", "phi2"),
    "second_order_upwind": ("us", "vs", "ws"),
}


def golden_inputs(case: M3Case, bubbles_only: bool = False):
    """inputs of the single-operation fixtures: the two bubbles, a seeded
    velocity (old level 0.9 of it), and the bookkeeping of step 5 (so
    SOLVE_LS reinitialises); the density, viscosity, normals and curvature
    that SECOND_ORDER_UPWIND reads are AMURHO / LS_CCD of the same phi,
    computed by the oracle (itself pinned by the amurho / ls_ccd entries)."""
    from oracle.m3 import M3Oracle
    inp = bubbles(case)
    if bubbles_only:
        return inp, {}
    inp.update(smooth_velocity(case, 3))
    state = dict(b1_im=5.0, b2_im=5.2, iter=5, rec_mass=1)
    o = M3Oracle(case, **state)
    o.set("phi", inp["phi"])
    o.call("amurho")
    o.call("ls_ccd")
    for k in ("rho", "amu", "normal_x", "normal_y", "normal_z", "curv"):
        inp[k] = o.get(k)
    inp["rho_old"], inp["amu_old"] = inp["rho"], inp["amu"]
    for k in ("u", "v", "w"):
        inp[k + "_old"] = 0.9 * inp[k]
    return inp, state


# oracle array name -> device field name (the reference's arrays on the
# context fields of the MPLS_RisingBubble step, include/fcfd.h)
M3_TO_DEV = {
    "phi": "phi", "phi1": "phi1", "phi2": "phi2", "phi_old": "phi_old", "u": "u", "v": "v", "w": "w",
    "u_old": "u_old", "v_old": "v_old", "w_old": "w_old", "uh": "nu", "vh": "nv", "wh": "nw",
    "us": "src_u", "vs": "src_v", "ws": "src_w", "us_old": "src_u_old", "vs_old": "src_v_old",
    "ws_old": "src_w_old", "rho": "rho", "rho_old": "rho_old", "amu": "mu", "amu_old": "mu_old", "p": "p",
    "rhs": "ppe_src", "heavy": "heavy", "delta": "delta", "grad": "grad", "sgn": "sign", "normal_x": "nx",
    "normal_y": "ny", "normal_z": "nz", "curv": "curv",
}


def device_case(case: M3Case):
    """the fcfd Case of an M3Case (3D, no-slip walls, the reference's
    INTERFACE_WIDTH = 1.5 dx, RDT = 0.5 dx)."""
    from fortrancfd_amd.types import Case
    return Case(dim=3, nx=case.nx, ny=case.ny, nz=case.nz, xend=case.nx * case.dx, yend=case.ny * case.dx,
                zend=case.nz * case.dx, cfl=case.cfl, cfl_r=0.5, ls_w=1.5, rho_12=case.ratio_rho,
                mu_12=case.ratio_amu, re=case.re, we=case.we, fr=case.fr, btn_sf=float(case.st_force),
                btn_g=float(case.g_force), gvec=(0.0, 0.0, -1.0), ppe_tol=case.ppe_tol,
                ppe_precision=case.ppe_precision, name="m3")
