"""The reference's time loop with its outputs (`Two phase flows/newOpenMP/
2D/src/main.f90:8-45`, `src/setup/problem_init.f90:79-90`): problem_init,
the plot snapshot and the `_mesh.vtk` header, then per step the whole
main-loop body on the device (`Solver.step`: time and iteration counters,
switch, interface solver, NS solver, ls_mv), a plot snapshot whenever
|time - pid * t2p| <= dt (plot.f90:10), one `_MVloss.plt` row (output.f90:8),
until time > t2s."""
from __future__ import annotations

import os

import numpy as np

from .io import MVLossLog, write_vts
from .solver import Solver
from .types import Case


def write_mesh_vtk(path: str, case: Case) -> None:
    """`manager_plot` (`lib/data/tree.f90:254-276`): a legacy-VTK
    STRUCTURED_POINTS header for the 2D cell-centre lattice."""
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nvtk TEST\nASCII\nDATASET STRUCTURED_POINTS\n")
        fh.write("DIMENSIONS %6d%6d%6d\n" % (case.nx, case.ny, 1))
        fh.write("SPACING %15.4E%15.4E%15.4E\n" % (case.dx, case.dy, 0.0))
        fh.write("ORIGIN %15.4E%15.4E%15.4E\n" % (case.xstart + 0.5 * case.dx, case.ystart + 0.5 * case.dy, 0.0))
        fh.write("POINT_DATA %12d\n" % (case.nx * case.ny))


class Plotter:
    """`plot`: numbered `<name>_NNN.vts` snapshots every t2p."""

    def __init__(self, solver: Solver, out_dir: str):
        self.s, self.out, self.pid = solver, out_dir, 0

    def __call__(self, time: float, force: bool = False) -> str:
        c = self.s.case
        if not force and abs(time - self.pid * c.t2p) > c.dt:
            return ""
        path = os.path.join(self.out, "%s_%03d.vts" % (c.name, self.pid))
        vel = [self.s.pull(n) for n in (("nu", "nv", "nw") if c.dim == 3 else ("nu", "nv"))]
        write_vts(path, c, self.s.pull("phi"), self.s.pull("p"), vel, c.num_of_plot)
        self.pid += 1
        return path


def run(case: Case, phi0: np.ndarray, out_dir: str, max_steps: int = 0) -> dict:
    """Run the reference main loop from the level set phi0 (ghosted) with
    the outputs written under out_dir; returns the last diagnostics.
    max_steps > 0 caps the loop (the reference runs until time > t2s)."""
    os.makedirs(out_dir, exist_ok=True)
    s = Solver(case)
    s.push("phi", phi0)
    s.problem_init(True)
    d = s.diag()
    plot = Plotter(s, out_dir)
    plot(0.0)
    if case.dim == 2:
        write_mesh_vtk(os.path.join(out_dir, "%s_mesh.vtk" % case.name), case)
    log = MVLossLog(os.path.join(out_dir, "%s_MVloss.plt" % case.name), d["mass"], d["vol"])
    n = 0
    while True:
        s.step()
        d = s.diag()
        time = d["iter"] * case.dt
        plot(time)
        log.row(time, d["mass"], d["vol"])
        n += 1
        if time > case.t2s or (max_steps and n >= max_steps):
            break
    return d
