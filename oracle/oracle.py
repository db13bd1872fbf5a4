"""ctypes binding to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It mirrors the C-ABI of the product library field-for-field so the
parity tests read the same on both sides.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from fortrancfd_amd.types import FIELD_ID, Case, FcfdDiag  # noqa: E402

LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

EXTRA_ID = {
    "rho_old": 100, "mu_old": 101, "delta_old": 102, "curv_old": 103, "nx_old": 104,
    "ny_old": 105, "nz_old": 106, "coe_r": 107, "coe_l": 108, "coe_f": 109, "coe_b": 110,
    "coe_u": 111, "coe_d": 112, "coe_c": 113,
}

_lib = None


def build(with_ref: bool = False) -> None:
    env = dict(os.environ)
    if not with_ref:
        env["FCFD_SKIP_REF"] = "1"
    subprocess.run(["bash", os.path.join(_HERE, "build.sh")], check=True, env=env)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_set_imass.argtypes = [C.c_void_p, C.c_double]
        L.orc_get_diag.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_band_max.argtypes = [C.c_long]
        for nm in ("orc_ppe_sor_solver", "orc_ppe_exact_solver", "orc_ppe_solve_rhs"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_double]
        for nm in ("orc_level_set_rk3_redis", "orc_ns_solver", "orc_step", "orc_problem_init", "orc_bc"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def field_id(name: str) -> int:
    return EXTRA_ID[name] if name in EXTRA_ID else FIELD_ID[name]


class Oracle:
    """One oracle state (the reference's `p`), for one grid."""

    def __init__(self, case: Case):
        self.case = case
        self._g = case.grid_struct()
        self._p = case.param_struct()
        self.h = lib().orc_create(C.byref(self._g), C.byref(self._p))
        if not self.h:
            raise ValueError("oracle: bad grid")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @staticmethod
    def threads(n: int) -> int:
        return lib().orc_set_threads(n)

    def set(self, name: str, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.case.size, (name, a.size, self.case.size)
        rc = lib().orc_set(self.h, field_id(name), a.ctypes.data)
        assert rc == 0, name

    def get(self, name: str) -> np.ndarray:
        a = np.empty(self.case.size, dtype=np.float64)
        rc = lib().orc_get(self.h, field_id(name), a.ctypes.data)
        assert rc == 0, name
        return a.reshape(self.case.shape)

    def set_time(self, time: float, it: int) -> None:
        lib().orc_set_time(self.h, time, it)

    def set_imass(self, m: float) -> None:
        lib().orc_set_imass(self.h, m)

    def diag(self) -> dict:
        d = FcfdDiag()
        lib().orc_get_diag(self.h, C.byref(d))
        return d.as_dict()

    def call(self, name: str, *args) -> int:
        fn = getattr(lib(), "orc_" + name)
        cargs = [C.c_int(a) if isinstance(a, (int, np.integer)) else C.c_double(a) for a in args]
        return fn(C.c_void_p(self.h), *cargs)
