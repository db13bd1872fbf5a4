"""Closed-form pins of the oracle's level-set sources, for what the
reference harness does not reach on its own: the UCCD advection source
s = -(u phi_x + v phi_y + w phi_z) composes the bitwise-pinned UCCD line
solver with the cell velocities (the reference keeps this form only as a
commented alternative, `3D+IBM/src/interface_solver/level_set_rk3_solver.f90:
155-157`).  A linear phi is differentiated exactly by every CCD row, so the
source must equal -(u a + v b + w c) cell by cell; a swapped velocity
component or a sign error fails it.  The WENO5 flux-split source is checked
the same way for a uniform velocity."""
from __future__ import annotations

import numpy as np
import pytest

from cases import case2d, case3d, seeded
from oracle.oracle import Oracle

GRAD = (0.7, -1.3, 0.45)


def _linear_phi(case):
    C = case.coords()
    return sum(GRAD[d] * C[d] for d in range(case.dim))


def _source(case, scheme, vel):
    case.ls_scheme = scheme
    Oracle.threads(1)
    o = Oracle(case)
    o.set("phi", _linear_phi(case))
    for k, v in vel.items():
        o.set(k, v)
    o.call("switch")            # nvel%old <- nvel%now
    o.call("level_set_source")
    return o.get("grad")


@pytest.mark.parametrize("mk", [lambda: case2d(24), lambda: case3d(12)])
def test_uccd_source_of_linear_phi(mk):
    case = mk()
    names = ["nu", "nv"] + (["nw"] if case.dim == 3 else [])
    vel = {k: seeded(case, 30 + i) for i, k in enumerate(names)}   # signs of both kinds
    L = case.interior(_source(case, 1, vel))
    want = -sum(GRAD[i] * case.interior(vel[k]) for i, k in enumerate(names))
    assert np.abs(L - want).max() <= 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("mk", [lambda: case2d(24), lambda: case3d(12)])
def test_weno_source_of_linear_phi_uniform_velocity(mk):
    case = mk()
    names = ["nu", "nv"] + (["nw"] if case.dim == 3 else [])
    speeds = (0.8, -0.35, 0.6)
    vel = {k: np.full(case.shape, speeds[i]) for i, k in enumerate(names)}
    L = _source(case, 0, vel)
    want = -sum(GRAD[i] * speeds[i] for i in range(case.dim))
    # the flux split is ghost-copied (bc), so only cells >= 3 from every wall
    # see linear data on their whole 7-point stencil
    Li = case.interior(L)
    core = Li[tuple(slice(3, s - 3) for s in Li.shape)]
    assert np.abs(core - want).max() <= 1e-11 * abs(want)


def test_uccd_derivative_converges():
    """phi = sin(2 pi x) advected by u = 1: the UCCD source converges at high
    order as the grid is refined (error ratio > 2^3.5 from 32 to 64 cells)."""
    errs = []
    for n in (32, 64):
        case = case2d(n)
        X = case.coords()[0]
        phi = np.sin(2 * np.pi * X)
        case.ls_scheme = 1
        o = Oracle(case)
        o.set("phi", phi)
        o.set("nu", np.ones(case.shape))
        o.set("nv", np.zeros(case.shape))
        o.call("switch")
        o.call("level_set_source")
        L = case.interior(o.get("grad"))
        want = -2 * np.pi * np.cos(2 * np.pi * case.interior(X))
        errs.append(np.abs(L - want).max())
    assert errs[0] / errs[1] > 2 ** 3.5, errs
