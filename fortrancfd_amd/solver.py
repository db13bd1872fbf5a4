"""Host-side mirror of the reference's `type(manager) :: p` over libfcfd.

`Solver` owns one libfcfd context (device-resident fields in the reference
layout) and exposes the reference's hot-path subroutines under their own
names: level_set_rk3_solver, level_set_rk3_redis(btn),
mass_preserving_level_set, ls_funs, rho_mu, ls_mv, surface_norms, curv,
ppe_sor_init, ppe_sor_solver, ns_momentum_correction, ns_ab_predictor,
node_vel, ns_solver, switch, interface_solver, plus step() for one pass of
the main loop (`Two phase flows/newOpenMP/3D+IBM/src/main.f90:15-27`).

Errors: the reference `stop`s on a diverging pressure solve
(`.../ns_solver/ns_ppe_solver.f90:352-355`); here every non-zero status of
the C ABI raises FcfdError (NOT_CONVERGED included unless allow_unconverged).
There is no CPU fallback: without the built library or a HIP device the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .types import FIELD_ID, PPE_FP64, STATUS, Case, FcfdCase, FcfdDiag, FcfdM3State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfcfd.so")

_lib = None


class FcfdError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed: status {code} ({STATUS.get(code, '?')})")
        self.code = code


def lib():
    """Load libfcfd.so (built by fortrancfd_amd/build.py); raises when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
    sig = {
        "fcfd_create": [vp, vp, C.POINTER(vp)],
        "fcfd_destroy": [vp],
        "fcfd_set_params": [vp, vp],
        "fcfd_push": [vp, i32, vp], "fcfd_pull": [vp, i32, vp],
        "fcfd_push_device": [vp, i32, vp], "fcfd_pull_device": [vp, i32, vp],
        "fcfd_sync": [vp], "fcfd_get_diag": [vp, vp],
        "fcfd_set_time": [vp, dbl, i32], "fcfd_set_imass": [vp, dbl],
        "fcfd_bc": [vp, i32], "fcfd_velbc": [vp],
        "fcfd_ls_funs": [vp], "fcfd_rho_mu": [vp], "fcfd_ls_mv": [vp, vp, vp],
        "fcfd_level_set_rk3_solver": [vp], "fcfd_level_set_rk3_redis": [vp, i32, vp],
        "fcfd_mass_preserving_level_set": [vp], "fcfd_surface_norms": [vp], "fcfd_curv": [vp], "fcfd_uccd": [vp],
        "fcfd_ls_adv_adjoint_step": [vp, vp, vp, vp, vp, vp],
        "fcfd_ls_adv_adjoint_step_f32": [vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "fcfd_set_velocity_scale": [vp, dbl], "fcfd_level_set_rk3_solver_f32": [vp, vp, vp, vp, vp],
        "fcfd_ppe_sor_init": [vp], "fcfd_ppe_sor_solver": [vp, dbl, i32, vp, vp],
        "fcfd_ppe_solve_rhs": [vp, dbl, i32, vp, vp],
        "fcfd_ns_momentum_correction": [vp], "fcfd_ns_ab_predictor": [vp],
        "fcfd_node_vel": [vp], "fcfd_ns_check_convergence_div": [vp], "fcfd_ns_solver": [vp],
        "fcfd_switch": [vp], "fcfd_interface_solver": [vp], "fcfd_step": [vp],
        "fcfd_problem_init": [vp, i32],
        "fcfd_create_batch": [vp, vp, i32, vp, C.POINTER(vp)], "fcfd_set_cases": [vp, vp],
        "fcfd_get_cases": [vp, vp], "fcfd_get_case_diag": [vp, vp, vp, vp], "fcfd_batch_size": [vp],
        "fcfd_m3_get_state": [vp, vp], "fcfd_m3_set_state": [vp, vp], "fcfd_m3_amurho": [vp],
        "fcfd_m3_ls_ccd": [vp], "fcfd_m3_ls_rk3_solver": [vp], "fcfd_m3_ls_redistance": [vp],
        "fcfd_m3_mpls": [vp, dbl], "fcfd_m3_solve_ls": [vp], "fcfd_m3_solve_ns": [vp], "fcfd_m3_find_mass": [vp],
        "fcfd_m3_store_data": [vp], "fcfd_m3_setup": [vp], "fcfd_m3_step": [vp],
        "fcfd_ls_mv_noalloc": [vp, vp, i32, vp] + [vp] * 7 + [vp],
        "fcfd_set_tile": [vp, i32, i32, i32, i32], "fcfd_get_tile": [vp, i32, vp, vp, vp],
        "fcfd_blocks_create": [vp, vp, i32, i32, i32, i32, i32, i32, C.POINTER(vp)],
        "fcfd_blocks_destroy": [vp], "fcfd_blocks_range": [vp, i32, i32, i32, vp, vp],
        "fcfd_blocks_scatter": [vp, vp, vp, vp], "fcfd_blocks_gather": [vp, vp, vp, vp],
        "fcfd_blocks_bc": [vp, vp, vp], "fcfd_blocks_sync": [vp, vp, vp],
        "fcfd_blocks_check_per": [vp, vp, vp, vp],
        "fcfd_blocks_level_set_rk3_solver": [vp, vp, vp, vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.fcfd_field_size.argtypes = [vp]
    L.fcfd_field_size.restype = C.c_int64
    L.fcfd_blocks_stack_size.argtypes = [vp]
    L.fcfd_blocks_stack_size.restype = C.c_int64
    L.fcfd_prof_nregions.argtypes = []
    L.fcfd_prof_nregions.restype = C.c_int
    L.fcfd_prof_region_name.argtypes = [C.c_int]
    L.fcfd_prof_region_name.restype = C.c_char_p
    L.fcfd_prof_enable.argtypes = [vp, i32]
    L.fcfd_prof_enable.restype = C.c_int
    L.fcfd_prof_read.argtypes = [vp, vp, vp, i32]
    L.fcfd_prof_read.restype = C.c_int
    _lib = L
    return L


class Solver:
    """One grid on the current HIP device (the reference's `p`)."""

    def __init__(self, case: Case, allow_unconverged: bool = False):
        self.case = case
        self.allow_unconverged = allow_unconverged
        self._grid = case.grid_struct()
        self._params = case.param_struct()
        h = C.c_void_p()
        self._check("fcfd_create", lib().fcfd_create(C.byref(self._grid), C.byref(self._params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fcfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, fn: str, rc: int):
        if rc == 3 and getattr(self, "allow_unconverged", False):
            return rc
        if rc != 0:
            raise FcfdError(fn, rc)
        return rc

    def _call(self, name: str, *args):
        return self._check(name, getattr(lib(), name)(self.h, *args))

    # ---- parameters / fields -------------------------------------------
    def set_params(self, **kw):
        if "imass" not in kw:
            # keep the MPLS target that problem_init measured on the device
            self.case.imass = self.diag()["imass"]
        for k, v in kw.items():
            setattr(self.case, k, v)
        self._params = self.case.param_struct()
        self._call("fcfd_set_params", C.byref(self._params))

    def push(self, name: str, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != self.case.size:
            raise ValueError(f"{name}: {a.size} values, field holds {self.case.size}")
        self._call("fcfd_push", FIELD_ID[name], a.ctypes.data)

    def pull(self, name: str) -> np.ndarray:
        a = np.empty(self.case.size, dtype=np.float64)
        self._call("fcfd_pull", FIELD_ID[name], a.ctypes.data)
        return a.reshape(self.case.shape)

    def push_device(self, name: str, ptr: int) -> None:
        self._call("fcfd_push_device", FIELD_ID[name], C.c_void_p(ptr))

    def pull_device(self, name: str, ptr: int) -> None:
        self._call("fcfd_pull_device", FIELD_ID[name], C.c_void_p(ptr))

    def sync(self) -> None:
        self._call("fcfd_sync")

    def diag(self) -> dict:
        d = FcfdDiag()
        self._call("fcfd_get_diag", C.byref(d))
        return d.as_dict()

    # ---- built-in kernel timing ----------------------------------------
    def prof_enable(self, on: bool = True) -> None:
        self._call("fcfd_prof_enable", 1 if on else 0)

    def prof_read(self) -> dict:
        """{region: (milliseconds, launches)} since prof_enable."""
        n = lib().fcfd_prof_nregions()
        ms = np.zeros(n, dtype=np.float64)
        cnt = np.zeros(n, dtype=np.int64)
        self._call("fcfd_prof_read", ms.ctypes.data, cnt.ctypes.data, n)
        return {lib().fcfd_prof_region_name(r).decode(): (float(ms[r]), int(cnt[r])) for r in range(n)}

    def set_tile(self, kernel: int, tx: int, ty: int, slab: int = 0) -> None:
        """thread-block tile and XCD slab order of one stencil kernel
        (types.TILE_KERNELS)"""
        self._call("fcfd_set_tile", int(kernel), int(tx), int(ty), int(slab))

    def get_tile(self, kernel: int):
        """(tx, ty, slab)"""
        tx, ty, sl = C.c_int(), C.c_int(), C.c_int()
        self._call("fcfd_get_tile", int(kernel), C.byref(tx), C.byref(ty), C.byref(sl))
        return tx.value, ty.value, sl.value

    def set_time(self, time: float, it: int) -> None:
        self._call("fcfd_set_time", C.c_double(time), it)

    def set_imass(self, m: float) -> None:
        self.case.imass = m
        self._call("fcfd_set_imass", C.c_double(m))

    # ---- reference subroutines -------------------------------------------
    def bc(self, name: str) -> None:
        self._call("fcfd_bc", FIELD_ID[name])

    def velbc(self) -> None:
        self._call("fcfd_velbc")

    def ls_funs(self) -> None:
        self._call("fcfd_ls_funs")

    def rho_mu(self) -> None:
        self._call("fcfd_rho_mu")

    def ls_mv(self):
        m, v = C.c_double(), C.c_double()
        self._call("fcfd_ls_mv", C.byref(m), C.byref(v))
        return m.value, v.value

    def level_set_rk3_solver(self) -> None:
        self._call("fcfd_level_set_rk3_solver")

    def set_velocity_scale(self, s: float) -> None:
        """The WENO5 advection (and its FP32 form and adjoint) uses s (NU, NV, NW)."""
        self._call("fcfd_set_velocity_scale", C.c_double(s))

    def level_set_rk3_solver_f32(self, phi_ptr: int, nu_ptr: int, nv_ptr: int, nw_ptr: int = 0) -> None:
        """One FP32 SSP-RK3 WENO5 step of the device array phi (ghosted
        layout, float32) with the float32 cell velocity (device pointers)."""
        self._call("fcfd_level_set_rk3_solver_f32", C.c_void_p(phi_ptr), C.c_void_p(nu_ptr), C.c_void_p(nv_ptr),
                   C.c_void_p(nw_ptr or None))

    def ls_adv_adjoint_step_f32(self, phi_ptr: int, lam_ptr: int, ub_ptr: int, vb_ptr: int, wb_ptr: int,
                                nu_ptr: int, nv_ptr: int, nw_ptr: int = 0) -> None:
        """FP32 adjoint of one level_set_rk3_solver_f32 step at phi (float32
        device arrays) with the float32 cell velocity nu, nv, nw times the
        current scale; u/v/w bars (0: not wanted) accumulate."""
        self._call("fcfd_ls_adv_adjoint_step_f32", C.c_void_p(phi_ptr), C.c_void_p(lam_ptr),
                   C.c_void_p(ub_ptr or None), C.c_void_p(vb_ptr or None), C.c_void_p(wb_ptr or None),
                   C.c_void_p(nu_ptr), C.c_void_p(nv_ptr), C.c_void_p(nw_ptr or None))

    def ls_adv_adjoint_step(self, phi_ptr: int, lam_ptr: int, ub_ptr: int = 0, vb_ptr: int = 0,
                            wb_ptr: int = 0) -> None:
        """Adjoint of one level_set_rk3_solver step at the level set phi
        (device pointer, ghosted layout) with the current s (NU, NV, NW): lambda
        (device) goes from dJ/dphi after the step to dJ/dphi before it; the
        velocity adjoint is added into the u/v/w bar arrays when given."""
        self._call("fcfd_ls_adv_adjoint_step", C.c_void_p(phi_ptr), C.c_void_p(lam_ptr),
                   C.c_void_p(ub_ptr or None), C.c_void_p(vb_ptr or None), C.c_void_p(wb_ptr or None))

    def level_set_rk3_redis(self, btn: int) -> int:
        it = C.c_int()
        self._call("fcfd_level_set_rk3_redis", btn, C.byref(it))
        return it.value

    def mass_preserving_level_set(self) -> None:
        self._call("fcfd_mass_preserving_level_set")

    def surface_norms(self) -> None:
        self._call("fcfd_surface_norms")

    def uccd(self) -> None:
        """root_solve("uccd") of phi with nvel%old: phi_x -> nx/ny/nz, fxx -> nxx/nyy/nzz."""
        self._call("fcfd_uccd")

    def curv(self) -> None:
        self._call("fcfd_curv")

    def ppe_sor_init(self) -> None:
        self._call("fcfd_ppe_sor_init")

    def ppe_sor_solver(self, rel_tol: float = 1e-10, precision: int = PPE_FP64):
        cyc, rel = C.c_int(), C.c_double()
        self._call("fcfd_ppe_sor_solver", C.c_double(rel_tol), precision, C.byref(cyc), C.byref(rel))
        return cyc.value, rel.value

    def ppe_solve_rhs(self, rel_tol: float = 1e-10, precision: int = PPE_FP64):
        """A p = PPE_SRC with the current operator, from the current P."""
        cyc, rel = C.c_int(), C.c_double()
        self._call("fcfd_ppe_solve_rhs", C.c_double(rel_tol), precision, C.byref(cyc), C.byref(rel))
        return cyc.value, rel.value

    def ns_momentum_correction(self) -> None:
        self._call("fcfd_ns_momentum_correction")

    def ns_ab_predictor(self) -> None:
        self._call("fcfd_ns_ab_predictor")

    def node_vel(self) -> None:
        self._call("fcfd_node_vel")

    def ns_check_convergence_div(self) -> None:
        self._call("fcfd_ns_check_convergence_div")

    def ns_solver(self) -> None:
        self._call("fcfd_ns_solver")

    def switch(self) -> None:
        self._call("fcfd_switch")

    def interface_solver(self) -> None:
        self._call("fcfd_interface_solver")

    def step(self) -> None:
        self._call("fcfd_step")

    def problem_init(self, redis0: bool = True) -> None:
        self._call("fcfd_problem_init", 1 if redis0 else 0)


# parameters that may differ between the cases of one batched context
CASE_KEYS = ("rho_12", "mu_12", "re", "we", "fr")


class M3Solver(Solver):
    """The MPLS_RisingBubble step (`MPLS_RisingBubble/20190612/`) on one 3D
    grid: the reference's subroutines under their own names (AMURHO,
    LS_CCD, LS_RK3_SOLVER, LS_REDISTANCE, MPLS, SOLVE_LS, SOLVE_NS,
    FIND_MASS, STORE_DATA), setup() for the part of SETUP after the initial
    level sets are pushed and step() for one pass of its main loop."""

    def state(self) -> dict:
        st = FcfdM3State()
        self._call("fcfd_m3_get_state", C.byref(st))
        return st.as_dict()

    def set_state(self, **kw) -> None:
        st = FcfdM3State()
        self._call("fcfd_m3_get_state", C.byref(st))
        for k, v in kw.items():
            if k == "bubble_imass":
                st.bubble_imass[0], st.bubble_imass[1] = v
            else:
                setattr(st, k, v)
        self._call("fcfd_m3_set_state", C.byref(st))

    def amurho(self) -> None:
        self._call("fcfd_m3_amurho")

    def ls_ccd(self) -> None:
        self._call("fcfd_m3_ls_ccd")

    def ls_rk3_solver(self) -> None:
        self._call("fcfd_m3_ls_rk3_solver")

    def ls_redistance(self) -> None:
        self._call("fcfd_m3_ls_redistance")

    def mpls(self, im: float) -> None:
        self._call("fcfd_m3_mpls", C.c_double(im))

    def solve_ls(self) -> None:
        self._call("fcfd_m3_solve_ls")

    def solve_ns(self) -> None:
        self._call("fcfd_m3_solve_ns")

    def find_mass(self) -> None:
        self._call("fcfd_m3_find_mass")

    def store_data(self) -> None:
        self._call("fcfd_m3_store_data")

    def setup(self) -> None:
        self._call("fcfd_m3_setup")

    def step(self) -> None:
        self._call("fcfd_m3_step")


class BatchSolver(Solver):
    """Many independent 2D cases on one grid in one context
    (fcfd_create_batch): case q is slice q of every field, and every call
    advances all of them in the same launches.  The cases must agree on the
    grid and on every shared parameter (time steps, gravity, switches,
    boundary conditions, solver settings); density / viscosity ratios, Re,
    We and Fr may differ.  Fields are pushed and pulled as arrays of shape
    (ncases, ny + 6, nx + 6)."""

    def __init__(self, cases, allow_unconverged: bool = False):
        cases = list(cases)
        if not cases:
            raise ValueError("no cases")
        c0 = cases[0]
        if c0.dim != 2:
            raise ValueError("batched contexts hold 2D cases")
        shared = c0.param_struct()
        for c in cases[1:]:
            p = c.param_struct()
            for name, _ in shared._fields_:
                if name in CASE_KEYS or name == "imass":
                    continue
                a, b = getattr(shared, name), getattr(p, name)
                a = list(a) if hasattr(a, "__len__") else a
                b = list(b) if hasattr(b, "__len__") else b
                if a != b:
                    raise ValueError(f"cases differ in the shared parameter {name}: {a} vs {b}")
            if (c.nx, c.ny, c.xstart, c.xend, c.ystart, c.yend) != (c0.nx, c0.ny, c0.xstart, c0.xend, c0.ystart,
                                                                   c0.yend):
                raise ValueError("cases of one batch must share the grid")
        self.case = c0
        self.cases = cases
        self.n = len(cases)
        self.allow_unconverged = allow_unconverged
        self._grid = c0.grid_struct()
        self._params = shared
        self._cases = (FcfdCase * self.n)(*[self._case_struct(c) for c in cases])
        h = C.c_void_p()
        self._check("fcfd_create_batch", lib().fcfd_create_batch(C.byref(self._grid), C.byref(self._params), self.n,
                                                                 self._cases, C.byref(h)))
        self.h = h

    def set_params(self, **kw):
        """Change shared parameters of every case (per-case ones go through set_cases)."""
        bad = [k for k in kw if k in CASE_KEYS or k == "imass"]
        if bad:
            raise ValueError(f"per-case parameters {bad}: use set_cases")
        for c in self.cases:
            for k, v in kw.items():
                setattr(c, k, v)
        self._params = self.case.param_struct()
        self._call("fcfd_set_params", C.byref(self._params))

    def set_cases(self, **per_case) -> None:
        """Per-case parameters: keyword -> sequence of n values (rho_12, mu_12, re, we, fr)."""
        cur = (FcfdCase * self.n)()
        self._call("fcfd_get_cases", cur)
        for k, vals in per_case.items():
            if k not in CASE_KEYS:
                raise ValueError(f"{k} is not a per-case parameter")
            if len(vals) != self.n:
                raise ValueError(f"{k}: {len(vals)} values for {self.n} cases")
            for q, v in enumerate(vals):
                setattr(cur[q], k, float(v))
                setattr(self.cases[q], k, float(v))
        self._call("fcfd_set_cases", cur)

    @staticmethod
    def _case_struct(c: Case) -> FcfdCase:
        p = c.param_struct()
        return FcfdCase(**{k: getattr(p, k) for k in CASE_KEYS}, imass=p.imass)

    @property
    def shape(self):
        return (self.n,) + tuple(self.case.shape)

    def push(self, name: str, a) -> None:
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1)
        if a.size != self.n * self.case.size:
            raise ValueError(f"{name}: {a.size} values, the batch holds {self.n} x {self.case.size}")
        self._call("fcfd_push", FIELD_ID[name], a.ctypes.data)

    def pull(self, name: str) -> np.ndarray:
        a = np.empty(self.n * self.case.size, dtype=np.float64)
        self._call("fcfd_pull", FIELD_ID[name], a.ctypes.data)
        return a.reshape(self.shape)

    def case_diag(self) -> dict:
        """Per-case mass, volume (after the last step or problem_init) and MPLS target mass."""
        m, v, im = (np.zeros(self.n) for _ in range(3))
        self._call("fcfd_get_case_diag", m.ctypes.data, v.ctypes.data, im.ctypes.data)
        return {"mass": m, "vol": v, "imass": im}
