"""ctypes mirrors of include/fcfd_types.h and the reference configuration reader.

The configuration reader restates the reference's `manager_read` /
`manager_init` (`Two phase flows/newOpenMP/3D+IBM/lib/data/tree.f90:48-227`,
2D: `Two phase flows/newOpenMP/2D/lib/data/tree.f90:48-215`): the same
positional text format with `ooo` comment lines, the same derived quantities
(dt = CFL*dx, rdt = CFL_r*dx, ls_wid = w*dx, tolerances 0.1**x) and the same
non-dimensionalisation selector `how_to_paras`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# enum fcfd_field, in declaration order (include/fcfd_types.h)
FIELDS: List[str] = [
    "phi", "phi_old", "u", "v", "w", "u_old", "v_old", "w_old", "nu", "nv", "nw",
    "src_u", "src_v", "src_w", "src_u_old", "src_v_old", "src_w_old", "p",
    "heavy", "delta", "sign", "rho", "mu", "grad", "nx", "ny", "nz",
    "nxx", "nyy", "nzz", "nxy", "nxz", "nyz", "curv", "ppe_src",
    "phi1", "phi2", "rho_old", "mu_old",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
# enum fcfd_tile_kernel -> the built-in profiler region that times it
TILE_KERNELS = {0: "predictor", 1: "diff_source", 2: "ppe_src", 3: "projection", 4: "node_vel"}
Z_FIELDS = {"w", "w_old", "nw", "src_w", "src_w_old", "nz", "nzz", "nxz", "nyz"}

FCFD_OK = 0
STATUS = {
    0: "ok", 1: "bad argument", 2: "HIP error", 3: "not converged",
    4: "NaN/Inf", 5: "field not held", 6: "singular",
}

PPE_FP64 = 0
PPE_MIXED = 1


class FcfdGrid(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("ghc", C.c_int32), ("pad_", C.c_int32),
        ("xstart", C.c_double), ("xend", C.c_double), ("ystart", C.c_double),
        ("yend", C.c_double), ("zstart", C.c_double), ("zend", C.c_double),
    ]


class FcfdParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("rdt", C.c_double), ("ls_wid", C.c_double),
        ("rho_12", C.c_double), ("mu_12", C.c_double),
        ("re", C.c_double), ("we", C.c_double), ("fr", C.c_double),
        ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
        ("btn_sf", C.c_double), ("btn_g", C.c_double), ("imass", C.c_double),
        ("ppe_tol", C.c_double), ("p_w1", C.c_double), ("p_tol", C.c_double),
        ("ubc", C.c_int32 * 2), ("vbc", C.c_int32 * 2), ("wbc", C.c_int32 * 2),
        ("redis_interval", C.c_int32), ("picard_iters", C.c_int32),
        ("ppe_precision", C.c_int32), ("ppe_max_cycles", C.c_int32), ("ppe_krylov", C.c_int32),
        ("ppe_inner", C.c_int32), ("ls_scheme", C.c_int32), ("ppe_nu1", C.c_int32), ("ppe_nu2", C.c_int32),
        ("ppe_nu_coarse", C.c_int32), ("ppe_fused", C.c_int32), ("ppe_rho_half", C.c_int32),
    ]


class FcfdCase(C.Structure):
    """Per-case parameters of a batched 2D stack (fcfd_case)."""
    _fields_ = [
        ("rho_12", C.c_double), ("mu_12", C.c_double),
        ("re", C.c_double), ("we", C.c_double), ("fr", C.c_double), ("imass", C.c_double),
    ]


class FcfdM3State(C.Structure):
    """struct fcfd_m3_state: bookkeeping of the MPLS_RisingBubble step."""
    _fields_ = [
        ("bubble_imass", C.c_double * 2), ("imass", C.c_double), ("ivol", C.c_double),
        ("mass", C.c_double), ("vol", C.c_double), ("div", C.c_double), ("l2sum", C.c_double),
        ("picard_tol", C.c_double), ("rec_mass", C.c_int32), ("picard", C.c_int32),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["bubble_imass"] = (self.bubble_imass[0], self.bubble_imass[1])
        return d


class FcfdDiag(C.Structure):
    _fields_ = [
        ("mass", C.c_double), ("vol", C.c_double), ("imass", C.c_double),
        ("ppe_rel_resid", C.c_double), ("vel_div", C.c_double), ("vel_sdiv", C.c_double),
        ("mpls_lambda", C.c_double),
        ("ppe_cycles", C.c_int32), ("ppe_cycles_step", C.c_int32),
        ("redis_iters", C.c_int32), ("iter", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


@dataclass
class Case:
    """A grid plus its derived parameters; one per solver context.

    The fields are the reference's `p%glb` members after `manager_init`.
    """

    dim: int
    nx: int
    ny: int
    nz: int
    xstart: float = 0.0
    xend: float = 1.0
    ystart: float = 0.0
    yend: float = 1.0
    zstart: float = 0.0
    zend: float = 1.0
    ghc: int = 3
    cfl: float = 0.1            # dt / dx
    cfl_r: float = 0.1          # rdt / dx
    ls_w: float = 1.5           # ls_wid / dx
    rho_12: float = 1.0e-3
    mu_12: float = 1.0e-2
    re: float = 1.0e4
    we: float = 500.0
    fr: float = 1.0
    gvec: Sequence[float] = (0.0, -1.0, 0.0)
    btn_sf: float = 1.0
    btn_g: float = 1.0
    ubc: Sequence[int] = (1, 1)
    vbc: Sequence[int] = (1, 1)
    wbc: Sequence[int] = (1, 1)
    p_tol_exp: float = 7.0      # reference relaxation: max update < 0.1**p_tol_exp
    p_w1: float = 1.5
    ppe_tol: float = 1.0e-10
    redis_interval: int = 5
    picard_iters: int = 4
    ppe_precision: int = PPE_FP64
    ppe_max_cycles: int = 200
    ppe_krylov: int = 1
    ppe_inner: int = 1
    ls_scheme: int = 0          # 0: WENO5 flux split, 1: upwind CCD (FCFD_LS_*)
    ppe_nu1: int = 2            # RB-GS sweeps before / after the coarse correction
    ppe_nu2: int = 2
    ppe_nu_coarse: int = 6      # sweeps on levels of <= 1/32 of the finest cells
    ppe_fused: int = 0          # coarse FP32 sweeps: 0 auto, 1 / 2 per launch, 3 colour launches, 4 auto from level 1
    ppe_rho_half: int = 0       # 1: FP32 level-0 smoother reads the density in FP16 (3D: level-1 coefficients too)
    imass: float = 0.0
    name: str = "case"
    # output schedule of the reference driver (manager_read): plot pieces
    # per axis, stop time and plot interval; host-side only
    num_of_plot: int = 1
    t2s: float = 0.0
    t2p: float = 0.0

    # ---- derived (tree.f90:154-172) -----------------------------------
    @property
    def dx(self) -> float:
        return (self.xend - self.xstart) / self.nx

    @property
    def dy(self) -> float:
        return (self.yend - self.ystart) / self.ny

    @property
    def dz(self) -> float:
        return (self.zend - self.zstart) / self.nz if self.dim == 3 else 1.0

    @property
    def dt(self) -> float:
        return self.cfl * self.dx

    @property
    def rdt(self) -> float:
        return self.cfl_r * self.dx

    @property
    def ls_wid(self) -> float:
        return self.ls_w * self.dx

    @property
    def shape(self) -> tuple:
        """numpy shape of a ghosted field in C order (k, j, i)."""
        g = self.ghc
        if self.dim == 3:
            return (self.nz + 2 * g, self.ny + 2 * g, self.nx + 2 * g)
        return (self.ny + 2 * g, self.nx + 2 * g)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)

    def interior(self, a: np.ndarray) -> np.ndarray:
        g = self.ghc
        a = a.reshape(self.shape)
        if self.dim == 3:
            return a[g:-g, g:-g, g:-g]
        return a[g:-g, g:-g]

    def coords(self):
        """cell-centre coordinates over the ghosted box (tree.f90:158-176)."""
        g = self.ghc
        i = np.arange(-g, self.nx + g)
        j = np.arange(-g, self.ny + g)
        x = self.xstart + (i + 1 - 0.5) * self.dx
        y = self.ystart + (j + 1 - 0.5) * self.dy
        if self.dim == 3:
            k = np.arange(-g, self.nz + g)
            z = self.zstart + (k + 1 - 0.5) * self.dz
            Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
            return X, Y, Z
        Y, X = np.meshgrid(y, x, indexing="ij")
        return X, Y

    def grid_struct(self) -> FcfdGrid:
        return FcfdGrid(self.dim, self.nx, self.ny, self.nz if self.dim == 3 else 1, self.ghc, 0,
                        self.xstart, self.xend, self.ystart, self.yend,
                        self.zstart if self.dim == 3 else 0.0, self.zend if self.dim == 3 else 1.0)

    def gravity_unit(self):
        """unit gravity direction (tree.f90:215-221)."""
        gx, gy = float(self.gvec[0]), float(self.gvec[1])
        gz = float(self.gvec[2]) if self.dim == 3 and len(self.gvec) > 2 else 0.0
        if self.dim == 3:
            mag = math.sqrt(gx * gx + gy * gy + gz * gz)
        else:
            mag = math.sqrt(gx * gx + gy * gy)
        return gx / mag, gy / mag, (gz / mag if self.dim == 3 else 0.0)

    def param_struct(self) -> FcfdParams:
        gx, gy, gz = self.gravity_unit()
        p = FcfdParams()
        p.dt, p.rdt, p.ls_wid = self.dt, self.rdt, self.ls_wid
        p.rho_12, p.mu_12, p.re, p.we, p.fr = self.rho_12, self.mu_12, self.re, self.we, self.fr
        p.gx, p.gy, p.gz = gx, gy, gz
        p.btn_sf, p.btn_g, p.imass = self.btn_sf, self.btn_g, self.imass
        p.ppe_tol, p.p_w1, p.p_tol = self.ppe_tol, self.p_w1, 0.1 ** self.p_tol_exp
        for nm in ("ubc", "vbc", "wbc"):
            arr = getattr(p, nm)
            vals = getattr(self, nm)
            arr[0], arr[1] = int(vals[0]), int(vals[1])
        p.redis_interval, p.picard_iters = self.redis_interval, self.picard_iters
        p.ppe_precision, p.ppe_max_cycles = self.ppe_precision, self.ppe_max_cycles
        p.ppe_krylov, p.ppe_inner = self.ppe_krylov, self.ppe_inner
        p.ls_scheme = self.ls_scheme
        p.ppe_nu1, p.ppe_nu2, p.ppe_nu_coarse = self.ppe_nu1, self.ppe_nu2, self.ppe_nu_coarse
        p.ppe_fused = self.ppe_fused
        p.ppe_rho_half = self.ppe_rho_half
        return p

    def ug(self) -> int:
        """cells per unit length; the reference grid is uniform with
        node_x = ug*(xend-xstart) (tree.f90:146-148)."""
        ug = int(round(self.nx / (self.xend - self.xstart)))
        ok = int(ug * (self.xend - self.xstart)) == self.nx and int(ug * (self.yend - self.ystart)) == self.ny
        if self.dim == 3:
            ok = ok and int(ug * (self.zend - self.zstart)) == self.nz
        if not ok:
            raise ValueError("the reference input format needs a uniform grid (ug cells per unit length)")
        return ug

    # ---- reference-format input file (tree.f90:48-106) ------------------
    def to_reference_input(self, threads: int = 1) -> str:
        r = repr
        gx, gy = self.gvec[0], self.gvec[1]
        lines = [
            "ooo Method to capture interface ( LS / MPLS / CLSVOF )", "2",
            "ooo Number of threads", str(int(threads)),
            "ooo Plot split", str(self.num_of_plot),
            "ooo Name of the problem", self.name,
            "ooo Cells per unit length", str(self.ug()),
            "ooo Number of ghost cell", str(self.ghc),
            "ooo Range of domain in X direction", f"{r(self.xstart)} {r(self.xend)}",
            "ooo Range of domain in Y direction", f"{r(self.ystart)} {r(self.yend)}",
        ]
        if self.dim == 3:
            lines += ["ooo Range of domain in Z direction", f"{r(self.zstart)} {r(self.zend)}"]
        lines += [
            "ooo Time to stop / plot", f"{r(self.t2s or 1.0)} {r(self.t2p or 1.0)}",
            "ooo Crount number ( advection / redistance )", f"{r(self.cfl)} {r(self.cfl_r)}",
            "ooo PPE iteration ( tolerance / w1 / w2 / valve )", f"{r(self.p_tol_exp)} {r(self.p_w1)} 0.5 4.0",
            "ooo SRK6 iteration ( tolerance / w )", "10.0 0.5",
            "ooo Level set interface width", r(self.ls_w),
            "ooo Problem parameters ( Prescribed / L / L+U / L+T / U+T )", "1",
            "ooo Density / Viscosity of fluid 1 ( phi > 0 )", "1.0 1.0",
            "ooo Density / Viscosity of fluid 2 ( phi < 0 )", "1.0 1.0",
            "ooo Surface tension of fluid 1, 2 / Using ?", f"1.0 {r(self.btn_sf)}",
            "ooo Gravity / Using ?", f"1.0 {r(self.btn_g)}",
        ]
        if self.dim == 3:
            lines += ["ooo Gravity direction", f"{r(gx)} {r(gy)} {r(self.gvec[2])}"]
        else:
            lines += ["ooo Gravity direction", f"{r(gx)} {r(gy)}"]
        lines += [
            "ooo Characteristic Length, Velocity, Time", "1.0 1.0 1.0",
            "ooo Dimensionless parameters ( Re, Fr, We )", f"{r(self.re)} {r(self.fr)} {r(self.we)}",
            "ooo Density / Viscosity ratio of fluid 1, 2", f"{r(self.rho_12)} {r(self.mu_12)}",
            "ooo Velocity boundary condition in X direction", f"{self.ubc[0]} {self.ubc[1]}",
            "ooo Velocity boundary condition in Y direction", f"{self.vbc[0]} {self.vbc[1]}",
        ]
        if self.dim == 3:
            lines += ["ooo Velocity boundary condition in Z direction", f"{self.wbc[0]} {self.wbc[1]}"]
        return "\n".join(lines) + "\n"

    def fields(self) -> List[str]:
        return [f for f in FIELDS if self.dim == 3 or f not in Z_FIELDS]


def read_reference_input(path: str, dim: int) -> Case:
    """Parse a reference-format input file (`manager_read`, tree.f90:48-106)
    and derive the grid like `manager_init` (node_x = ug*(xend-xstart))."""
    with open(path) as fh:
        raw = fh.read().splitlines()
    vals = [ln.split() for k, ln in enumerate(raw) if k % 2 == 1]
    it = iter(vals)
    method = int(next(it)[0]); _threads = next(it); nplot = int(next(it)[0])
    name = next(it)[0]
    ug = int(float(next(it)[0]))
    ghc = int(next(it)[0])
    xs, xe = map(float, next(it)[:2])
    ys, ye = map(float, next(it)[:2])
    zs, ze = (0.0, 1.0)
    if dim == 3:
        zs, ze = map(float, next(it)[:2])
    t2s, t2p = map(float, next(it)[:2])
    cfl, cfl_r = map(float, next(it)[:2])
    ptol, pw1 = map(float, next(it)[:2])
    _srk = next(it)
    lsw = float(next(it)[0])
    how = int(next(it)[0])
    rho1, mu1 = map(float, next(it)[:2])
    rho2, mu2 = map(float, next(it)[:2])
    sigma, btn_sf = map(float, next(it)[:2])
    g, btn_g = map(float, next(it)[:2])
    gv = tuple(map(float, next(it)[: dim]))
    L, U, T = map(float, next(it)[:3])
    re, fr, we = map(float, next(it)[:3])
    r12, m12 = map(float, next(it)[:2])
    ubc = tuple(map(int, next(it)[:2]))
    vbc = tuple(map(int, next(it)[:2]))
    wbc = tuple(map(int, next(it)[:2])) if dim == 3 else (1, 1)
    if method != 2:
        raise ValueError("only the MPLS interface method (2) is on this path")
    # tree.f90:176-201
    if how == 2:
        U = math.sqrt(L * g)
    elif how == 3:
        pass
    elif how == 4:
        U = L / T
    elif how == 5:
        L = U * T
    elif how != 1:
        raise ValueError("Wrong parameter selector")
    if how > 1:
        m12 = mu2 / mu1
        r12 = rho2 / rho1
        re = rho1 * U * L / mu1
        we = rho1 * U ** 2.0 * L / sigma
        fr = U ** 2.0 / (g * L)
    nx = int(ug * (xe - xs)); ny = int(ug * (ye - ys)); nz = int(ug * (ze - zs)) if dim == 3 else 1
    gvec = gv + ((0.0,) if dim == 2 else ())
    return Case(dim=dim, nx=nx, ny=ny, nz=nz, xstart=xs, xend=xe, ystart=ys, yend=ye,
                zstart=zs, zend=ze, ghc=ghc, cfl=cfl, cfl_r=cfl_r, ls_w=lsw, rho_12=r12,
                mu_12=m12, re=re, we=we, fr=fr, gvec=gvec, btn_sf=btn_sf, btn_g=btn_g,
                ubc=ubc, vbc=vbc, wbc=wbc, p_tol_exp=ptol, p_w1=pw1, name=name,
                num_of_plot=nplot, t2s=t2s, t2p=t2p)


def rising_bubble(dim: int, nx: int, ny: int, nz: int = 1, *, lx: float = 1.0, ly: float = 2.0,
                  lz: float = 1.0, radius: float = 0.25, centre=(0.5, 0.5, 0.5), **kw) -> Case:
    """Rising-bubble case: gravity along -y (2D) / -z (3D), no-slip walls.

    The non-dimensional numbers default to the Hysing et al. TC2 fluid pair
    (density ratio 1000, viscosity ratio 100, We = 500, Fr = 1 with L = 1) with
    Re raised to 1e4 so the explicit AB2 viscous term is stable at dt = 0.1 dx.
    """
    if dim == 2:
        c = Case(dim=2, nx=nx, ny=ny, nz=1, xend=lx, yend=ly, gvec=(0.0, -1.0, 0.0), **kw)
    else:
        c = Case(dim=3, nx=nx, ny=ny, nz=nz, xend=lx, yend=ly, zend=lz, gvec=(0.0, 0.0, -1.0), **kw)
    c._bubble = (radius, tuple(centre))
    return c


def bubble_phi(case: Case, radius: float, centre) -> np.ndarray:
    """Signed distance to a sphere/circle, positive in the heavy fluid (phi > 0
    is fluid 1, `.../2D/src/interface_solver/others.f90:54`)."""
    if case.dim == 3:
        X, Y, Z = case.coords()
        d = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2)
    else:
        X, Y = case.coords()
        d = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2)
    return np.ascontiguousarray(d - radius, dtype=np.float64)


def bc_copy(case: Case, a: np.ndarray) -> np.ndarray:
    """Zero-gradient ghost fill (the reference's `bc`, bcs.f90:1-42): every
    ghost value equals the nearest interior value along each axis."""
    g = case.ghc
    a = a.reshape(case.shape)
    if case.dim == 3:
        a[:g] = a[g]; a[-g:] = a[-g - 1]
        a[:, :g] = a[:, g:g + 1]; a[:, -g:] = a[:, -g - 1:-g]
        a[:, :, :g] = a[:, :, g:g + 1]; a[:, :, -g:] = a[:, :, -g - 1:-g]
    else:
        a[:g] = a[g]; a[-g:] = a[-g - 1]
        a[:, :g] = a[:, g:g + 1]; a[:, -g:] = a[:, -g - 1:-g]
    return a
