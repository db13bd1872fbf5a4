"""Smoother-schedule autotuner for the pressure solve.

The multigrid's cost per solve depends on the smoother schedule (sweeps per
level, extra sweeps on the small levels, V-cycles per flexible-CG iteration)
and on how the FP32 coarse-level sweeps are tiled (two sweeps per launch
with a 4-cell halo, one with a 2-cell halo, or one colour per launch), and
the best choice moves with the grid and the number of cases in a batched
context: fewer, larger launches pay off where the device is underfilled.
`tune_ppe_schedule` times the solve of the solver's current pressure system
from p = 0 under each candidate and keeps the fastest one that converges.
The system (operator of the current density, rhs in PPE_SRC) is left as it
was and the pressure is restored, so tuning does not change the run.
"""
from __future__ import annotations

import time

from .solver import FcfdError

# (pre, post) sweeps, sweeps on the small levels, V-cycles per outer
# iteration, FP32 coarse tiling (0 auto, 1 / 2 sweeps per launch, 3 colour launches)
SCHEDULES = [
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=6, ppe_inner=1, ppe_fused=0),
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=6, ppe_inner=1, ppe_fused=1),
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=6, ppe_inner=1, ppe_fused=2),
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=6, ppe_inner=1, ppe_fused=3),
    dict(ppe_nu1=1, ppe_nu2=1, ppe_nu_coarse=4, ppe_inner=1, ppe_fused=0),
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=4, ppe_inner=1, ppe_fused=0),
    dict(ppe_nu1=2, ppe_nu2=2, ppe_nu_coarse=8, ppe_inner=1, ppe_fused=0),
    dict(ppe_nu1=3, ppe_nu2=3, ppe_nu_coarse=6, ppe_inner=1, ppe_fused=0),
    dict(ppe_nu1=1, ppe_nu2=1, ppe_nu_coarse=6, ppe_inner=2, ppe_fused=0),
]
KEYS = ("ppe_nu1", "ppe_nu2", "ppe_nu_coarse", "ppe_inner", "ppe_fused")


def tune_ppe_schedule(solver, candidates=SCHEDULES, reps: int = 2, tol: float = None, precision: int = None) -> dict:
    """Pick the fastest smoother schedule for `solver` (a Solver or
    BatchSolver whose pressure operator is set up, e.g. after a step) and
    apply it.  Returns {"best": schedule, "ms": {index: milliseconds}}."""
    import numpy as np
    case = solver.case
    tol = case.ppe_tol if tol is None else tol
    precision = case.ppe_precision if precision is None else precision
    keep = {k: getattr(case, k) for k in KEYS}
    p_saved = solver.pull("p")
    zeros = np.zeros_like(p_saved)
    times = {}
    for n, sched in enumerate(candidates):
        solver.set_params(**sched)
        best = None
        try:
            for rep in range(reps + 1):          # the first solve captures the cycle graphs
                solver.push("p", zeros)
                solver.sync()
                t = time.perf_counter()
                solver.ppe_solve_rhs(tol, precision)
                solver.sync()
                ms = (time.perf_counter() - t) * 1e3
                if rep > 0:
                    best = ms if best is None else min(best, ms)
        except FcfdError:
            best = None                          # did not converge: not a candidate
        if best is not None:
            times[n] = round(best, 3)
    solver.push("p", p_saved)
    if not times:
        solver.set_params(**keep)
        return {"best": keep, "ms": times}
    win = min(times, key=times.get)
    solver.set_params(**candidates[win])
    return {"best": dict(candidates[win]), "ms": times}
