"""Seeded input states shared by the parity tests (oracle vs reference,
HIP path vs oracle) and the golden-fixture generator.

Every state is a dict of full ghosted float64 arrays in the reference layout
(numpy C order on a (k, j, i) shape == Fortran column-major (i, j, k)).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from fortrancfd_amd.types import Case, bc_copy, bubble_phi  # noqa: E402


def case2d(n: int = 24, **kw) -> Case:
    """2D rising-bubble geometry on [0,1]x[0,2] with n x 2n cells."""
    kw.setdefault("name", "t2d")
    return Case(dim=2, nx=n, ny=2 * n, nz=1, xend=1.0, yend=2.0, gvec=(0.0, -1.0, 0.0), **kw)


def case3d(n: int = 12, **kw) -> Case:
    """3D rising-bubble geometry on [0,1]^2x[0,2] with n x n x 2n cells."""
    kw.setdefault("name", "t3d")
    return Case(dim=3, nx=n, ny=n, nz=2 * n, xend=1.0, yend=1.0, zend=2.0, gvec=(0.0, 0.0, -1.0), **kw)


def case3d_odd(**kw) -> Case:
    """3D box of 20 x 18 x 37 cells (dx = dy = dz = 0.05): no extent is a
    multiple of the 16-plane walks or of the 64 x 4 tiles of the kernels, so
    their remainder planes and partial tiles are exercised."""
    kw.setdefault("name", "t3d")
    return Case(dim=3, nx=20, ny=18, nz=37, xend=1.0, yend=0.9, zend=1.85, gvec=(0.0, 0.0, -1.0), **kw)


def _smooth(case: Case, rng, amp: float, nmodes: int = 3) -> np.ndarray:
    """a smooth seeded field: sum of a few low Fourier modes over the box."""
    if case.dim == 3:
        X, Y, Z = case.coords()
    else:
        X, Y = case.coords()
        Z = np.zeros_like(X)
    f = np.zeros_like(X)
    for _ in range(nmodes):
        kx, ky, kz = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        a = rng.uniform(-1, 1)
        f += a * np.sin(kx * np.pi * X + ph[0]) * np.sin(ky * np.pi * Y + ph[1]) * np.cos(kz * np.pi * Z + ph[2])
    return np.ascontiguousarray(amp * f)


def bubble(case: Case, perturb: float = 0.0, seed: int = 0) -> np.ndarray:
    """signed distance to the bubble (r = 0.25 at (0.5, 0.5[, 0.5])), optionally
    multiplied by a smooth positive factor so it is no longer a distance."""
    c = (0.5, 0.5, 0.5)
    phi = bubble_phi(case, 0.25, c)
    if perturb:
        rng = np.random.default_rng(seed)
        phi = phi * (1.0 + perturb * (1.0 + _smooth(case, rng, 0.5)))
    return bc_copy(case, phi)


def velocities(case: Case, seed: int, amp: float = 0.3) -> dict:
    rng = np.random.default_rng(seed)
    out = {"u": _smooth(case, rng, amp), "v": _smooth(case, rng, amp)}
    if case.dim == 3:
        out["w"] = _smooth(case, rng, amp)
    return out


def node_velocities(case: Case, seed: int, amp: float = 0.5) -> dict:
    rng = np.random.default_rng(seed + 1000)
    out = {"nu": _smooth(case, rng, amp), "nv": _smooth(case, rng, amp)}
    if case.dim == 3:
        out["nw"] = _smooth(case, rng, amp)
    return out


def seeded(case: Case, seed: int, amp: float = 1.0) -> np.ndarray:
    return _smooth(case, np.random.default_rng(seed + 7), amp)


def vel_names(case: Case):
    return ["u", "v", "w"] if case.dim == 3 else ["u", "v"]


def max_rel(a: np.ndarray, b: np.ndarray, interior_of: Case = None) -> float:
    """max |a - b| / max(max |b|, 1e-300), over the interior when a case is given."""
    if interior_of is not None:
        a = interior_of.interior(a)
        b = interior_of.interior(b)
    den = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / den
