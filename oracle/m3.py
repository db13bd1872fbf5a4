"""ctypes binding to the MPLS_RisingBubble oracle (m3_oracle.c in
oracle/_build/liboracle.so) and a runner for the compiled reference harness
(oracle/_ref/refm3).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only
tests/ and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from dataclasses import dataclass, field, fields

import numpy as np

from .oracle import REF_DIR, lib

FIELDS = ["phi", "phi1", "phi2", "phi_old", "u", "v", "w", "u_old", "v_old", "w_old", "uh", "vh", "wh",
          "us", "vs", "ws", "us_old", "vs_old", "ws_old", "u1", "v1", "w1", "rho", "rho_old", "amu",
          "amu_old", "p", "rhs", "heavy", "delta", "grad", "sgn", "normal_x", "normal_y", "normal_z",
          "curv", "s0"]
FID = {n: i for i, n in enumerate(FIELDS)}


class M3Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("st_force", C.c_int), ("g_force", C.c_int),
                ("press_max_it", C.c_int), ("iter", C.c_int), ("rec_mass", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("rdt", C.c_double), ("interface_width", C.c_double), ("ratio_rho", C.c_double),
                ("ratio_amu", C.c_double), ("re", C.c_double), ("we", C.c_double), ("fr", C.c_double),
                ("tol_p", C.c_double), ("tol_m", C.c_double), ("press_omega", C.c_double),
                ("b1_im", C.c_double), ("b2_im", C.c_double), ("imass_ls", C.c_double),
                ("ivol_ls", C.c_double), ("time", C.c_double)]


class M3Diag(C.Structure):
    _fields_ = [("mass_ls", C.c_double), ("vol_ls", C.c_double), ("div", C.c_double), ("l2sum", C.c_double),
                ("ppe_iters", C.c_int), ("picard", C.c_int)]


@dataclass
class M3Case:
    """The two-bubble problem of MPLS_RisingBubble/20190612/SETUP.F90 on a
    grid of nx x ny x nz cells of size dx (x, y, z from 0): its
    dimensionless numbers (Re = 67.27, We = 16, Fr = 1), density / viscosity
    ratios 0.001 / 0.01, dt = 0.005 dx, rdt = 0.5 dx, interface half-width
    1.5 dx, SOR weight 1.5 to max |dp| < 1e-7, Picard tolerance 1e-5."""
    nx: int
    ny: int
    nz: int
    dx: float
    ratio_rho: float = 0.001
    ratio_amu: float = 0.01
    re: float = 67.27
    we: float = 16.0
    fr: float = 1.0
    cfl: float = 0.005
    st_force: int = 1
    g_force: int = 1
    tol_p: float = 1.0e-7
    tol_m: float = 1.0e-5
    press_omega: float = 1.5
    press_max_it: int = 5000000
    ppe_tol: float = 1.0e-13     # device pressure solve of the product path
    ppe_precision: int = 0

    @property
    def shape(self):
        return (self.nz + 6, self.ny + 6, self.nx + 6)

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def dt(self):
        return self.cfl * self.dx

    def params(self, **state) -> M3Params:
        p = M3Params(nx=self.nx, ny=self.ny, nz=self.nz, st_force=self.st_force, g_force=self.g_force,
                     press_max_it=self.press_max_it, dx=self.dx, dy=self.dx, dz=self.dx, dt=self.dt,
                     rdt=0.5 * self.dx, interface_width=1.5 * self.dx, ratio_rho=self.ratio_rho,
                     ratio_amu=self.ratio_amu, re=self.re, we=self.we, fr=self.fr, tol_p=self.tol_p,
                     tol_m=self.tol_m, press_omega=self.press_omega)
        for k, v in state.items():
            setattr(p, k, v)
        return p

    def coords(self):
        """cell centres X(I) = (I - 0.5) DX over the ghosted box, (k, j, i) order"""
        def ax(n):
            return (np.arange(-2, n + 4) - 0.5) * self.dx
        Z, Y, X = np.meshgrid(ax(self.nz), ax(self.ny), ax(self.nx), indexing="ij")
        return X, Y, Z

    def interior(self, a):
        return np.asarray(a).reshape(self.shape)[3:-3, 3:-3, 3:-3]


def _lib():
    L = lib()
    if not getattr(L, "_m3_ready", False):
        L.m3_create.restype = C.c_void_p
        L.m3_create.argtypes = [C.c_void_p]
        L.m3_destroy.argtypes = [C.c_void_p]
        L.m3_set.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.m3_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.m3_set_params.argtypes = [C.c_void_p, C.c_void_p]
        L.m3_get_params.argtypes = [C.c_void_p, C.c_void_p]
        L.m3_get_diag.argtypes = [C.c_void_p, C.c_void_p]
        L.m3_mpls.argtypes = [C.c_void_p, C.c_double]
        for nm in ("m3_solve_ns", "m3_setup_tail", "m3_step"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int]
        L._m3_ready = True
    return L


class M3Oracle:
    """One oracle state of the MPLS_RisingBubble step."""

    def __init__(self, case: M3Case, **state):
        self.case = case
        self._p = case.params(**state)
        self.h = _lib().m3_create(C.byref(self._p))
        if not self.h:
            raise ValueError("m3 oracle: bad grid")

    def __del__(self):
        if getattr(self, "h", None):
            _lib().m3_destroy(self.h)
            self.h = None

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.case.size
        assert _lib().m3_set(self.h, FID[name], a.ctypes.data) == 0

    def get(self, name):
        a = np.empty(self.case.size)
        assert _lib().m3_get(self.h, FID[name], a.ctypes.data) == 0
        return a.reshape(self.case.shape)

    def params(self) -> dict:
        p = M3Params()
        _lib().m3_get_params(self.h, C.byref(p))
        return {k: getattr(p, k) for k, _ in M3Params._fields_}

    def set_state(self, **kw):
        p = M3Params()
        _lib().m3_get_params(self.h, C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        _lib().m3_set_params(self.h, C.byref(p))

    def diag(self) -> dict:
        d = M3Diag()
        _lib().m3_get_diag(self.h, C.byref(d))
        return {k: getattr(d, k) for k, _ in M3Diag._fields_}

    def call(self, name, *args):
        fn = getattr(_lib(), "m3_" + name)
        return fn(C.c_void_p(self.h), *args)


# ---------------------------------------------------------------- harness
HARNESS = os.path.join(REF_DIR, "refm3")


def harness_available() -> bool:
    return os.access(HARNESS, os.X_OK)


def run_reference(case: M3Case, op: str, inputs: dict, timeout: float = 900.0, **state) -> dict:
    """Run one op of the compiled reference harness (single thread) on the
    given fields; returns every field and the scalars."""
    p = case.params(**state)
    with tempfile.TemporaryDirectory(prefix="fcfdm3_") as td:
        ind, outd = os.path.join(td, "in"), os.path.join(td, "out")
        os.makedirs(ind)
        os.makedirs(outd)
        with open(os.path.join(ind, "params.txt"), "w") as fh:
            fh.write(f"{p.nx} {p.ny} {p.nz}\n")
            fh.write(f"0.0 {p.nx * p.dx!r} 0.0 {p.ny * p.dy!r} 0.0 {p.nz * p.dz!r}\n")
            fh.write(f"{p.dx!r} {p.dy!r} {p.dz!r}\n")
            fh.write(f"{p.dt!r} {p.rdt!r} {p.interface_width!r}\n")
            fh.write(f"{p.ratio_rho!r} {p.ratio_amu!r} {p.re!r} {p.we!r} {p.fr!r}\n")
            fh.write(f"{p.st_force} {p.g_force} 0\n")
            fh.write(f"{p.tol_p!r} {p.tol_m!r} {p.press_omega!r} {p.press_max_it}\n")
            fh.write(f"{p.b1_im!r} {p.b2_im!r} {p.imass_ls!r} {p.ivol_ls!r} {p.time!r} {p.iter} {p.rec_mass}\n")
        for k, a in inputs.items():
            if k not in FID:
                raise KeyError(k)
            np.ascontiguousarray(a, dtype="<f8").reshape(-1).tofile(os.path.join(ind, k))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([HARNESS, op, "in", "out"], cwd=td, env=env, capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"m3 harness failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        out = {k: np.fromfile(os.path.join(outd, k), dtype="<f8").reshape(case.shape) for k in FIELDS}
        v = open(os.path.join(outd, "scalars.txt")).read().split()
        keys = ["mass_ls", "vol_ls", "imass_ls", "ivol_ls", "b1_im", "b2_im", "time", "div", "l2sum"]
        out["_scalars"] = {k: float(x) for k, x in zip(keys, v[:9])}
        out["_scalars"]["iter"] = int(v[9])
        out["_scalars"]["rec_mass"] = int(v[10])
        out["_stdout"] = r.stdout
        return out
