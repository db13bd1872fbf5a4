"""Run the compiled Fortran reference harness (oracle/_ref/ref2d|ref3d).

TEST INFRASTRUCTURE ONLY: produces golden vectors from the reference itself so
the CPU oracle can be pinned.  Field files are raw float64 full ghosted arrays
in the reference's column-major layout (i fastest), which is numpy C order on
a (k, j, i) shape.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from .oracle import REF_DIR

HARNESS_FIELDS_3D = [
    "phi", "phi_old", "u", "v", "w", "u_old", "v_old", "w_old", "nu", "nv", "nw",
    "src_u", "src_v", "src_w", "src_u_old", "src_v_old", "src_w_old", "p", "heavy", "delta",
    "sign", "rho", "mu", "grad", "nx", "ny", "nz", "nxx", "nyy", "nzz", "nxy", "nxz", "nyz",
    "curv", "ppe_src", "rho_old", "mu_old", "delta_old", "curv_old", "nx_old", "ny_old",
    "nz_old", "coe_r", "coe_l", "coe_f", "coe_b", "coe_u", "coe_d", "coe_c",
]
_Z = {"w", "w_old", "nw", "src_w", "src_w_old", "nz", "nzz", "nxz", "nyz", "nz_old", "coe_u", "coe_d"}
HARNESS_FIELDS_2D = [f for f in HARNESS_FIELDS_3D if f not in _Z]


def harness_path(dim: int) -> str:
    return os.path.join(REF_DIR, "ref3d" if dim == 3 else "ref2d")


def available(dim: int) -> bool:
    return os.access(harness_path(dim), os.X_OK)


def run(case, op: str, inputs: dict, imass: float = 0.0, time: float = 0.0, it: int = 0,
        timeout: float = 600.0, threads: int = 1) -> dict:
    names = HARNESS_FIELDS_3D if case.dim == 3 else HARNESS_FIELDS_2D
    with tempfile.TemporaryDirectory(prefix="fcfdref_") as td:
        os.makedirs(os.path.join(td, "in"))
        os.makedirs(os.path.join(td, "outf"))
        os.makedirs(os.path.join(td, "out"))
        with open(os.path.join(td, "case.txt"), "w") as fh:
            fh.write(case.to_reference_input(threads))
        for k, a in inputs.items():
            if k not in names:
                raise KeyError(k)
            np.ascontiguousarray(a, dtype="<f8").reshape(-1).tofile(os.path.join(td, "in", k))
        with open(os.path.join(td, "in", "scalars.txt"), "w") as fh:
            fh.write(f"{imass!r} {time!r} {it}\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(int(threads)))
        r = subprocess.run([harness_path(case.dim), op, "case.txt", "in", "outf"], cwd=td, env=env,
                           capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        out = {}
        for k in names:
            out[k] = np.fromfile(os.path.join(td, "outf", k), dtype="<f8").reshape(case.shape)
        with open(os.path.join(td, "outf", "scalars.txt")) as fh:
            v = fh.read().split()
        out["_mass"], out["_vol"], out["_imass"], out["_time"] = map(float, v[:4])
        out["_piter"] = int(v[4])
        out["_ppe_err"] = float(v[5])
        out["_stdout"] = r.stdout
        # files the reference wrote under ./out (_MVloss.plt, plot snapshots)
        od = os.path.join(td, "out")
        out["_files"] = {fn: open(os.path.join(od, fn), "rb").read() for fn in sorted(os.listdir(od))}
        return out


def run_steps_timed(case, inputs: dict, nsteps: int, imass: float, threads: int, timeout: float = 900.0) -> dict:
    """nsteps main-loop steps of the compiled reference in one process, timed
    inside it (system_clock around the loop only, no field I/O); returns
    {"seconds", "steps", "piter"}."""
    names = HARNESS_FIELDS_3D if case.dim == 3 else HARNESS_FIELDS_2D
    with tempfile.TemporaryDirectory(prefix="fcfdref_") as td:
        for d in ("in", "outf", "out"):
            os.makedirs(os.path.join(td, d))
        with open(os.path.join(td, "case.txt"), "w") as fh:
            fh.write(case.to_reference_input(threads))
        for k, a in inputs.items():
            if k not in names:
                raise KeyError(k)
            np.ascontiguousarray(a, dtype="<f8").reshape(-1).tofile(os.path.join(td, "in", k))
        with open(os.path.join(td, "in", "scalars.txt"), "w") as fh:
            fh.write(f"{imass!r} 0.0 0\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(int(threads)), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([harness_path(case.dim), "steps_timed", "case.txt", "in", "outf", str(int(nsteps))],
                           cwd=td, env=env, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        v = open(os.path.join(td, "outf", "timing.txt")).read().split()
        return {"seconds": float(v[0]), "steps": int(v[1]), "piter": int(v[2])}
