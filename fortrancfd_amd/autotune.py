"""Autotuners: the smoother schedule of the pressure solve and the
thread-block tiles of the stencil kernels.

The multigrid's cost per solve depends on the smoother schedule (sweeps per
level, extra sweeps on the small levels, V-cycles per flexible-CG iteration)
on how the FP32 coarse-level sweeps are tiled (two sweeps per launch with
a 4-cell halo, one with a 2-cell halo, or one colour per launch) and on
whether the FP32 level-0 sweeps read the density in FP16, and
the best choice moves with the grid and the number of cases in a batched
context: fewer, larger launches pay off where the device is underfilled.
`tune_ppe_schedule` times the solve of the solver's current pressure system
from p = 0 under each candidate and keeps the fastest one that converges.
The system (operator of the current density, rhs in PPE_SRC) is left as it
was and the pressure is restored, so tuning does not change the run.

`tune_tiles` picks the thread-block tile of each stencil kernel of the step
(predictor, viscous / body-force source, PPE source, projection, node
velocities).  A tile does not change any cell's arithmetic, so the results
are bit for bit the same under every tile, and the tuner times real steps:
each candidate runs whole steps with the built-in event regions on, every
kernel keeps the candidate whose mean launch time was least.
"""
from __future__ import annotations

import time

from .solver import FcfdError

# (pre, post) sweeps, sweeps on the small levels, V-cycles per outer
# iteration, FP32 coarse tiling (0 auto, 1 / 2 sweeps per launch, 3 colour
# launches, 4 auto from level 1 down), FP16 density in the FP32 level-0 sweeps
def _s(nu1, nu2, nuc, inner, fused, half=1):
    return dict(ppe_nu1=nu1, ppe_nu2=nu2, ppe_nu_coarse=nuc, ppe_inner=inner, ppe_fused=fused, ppe_rho_half=half)


SCHEDULES = [
    _s(2, 2, 6, 1, 0), _s(2, 2, 6, 1, 0, half=0), _s(2, 2, 6, 1, 1), _s(2, 2, 6, 1, 2), _s(2, 2, 6, 1, 3),
    _s(1, 1, 4, 1, 0), _s(2, 2, 4, 1, 0), _s(2, 2, 8, 1, 0), _s(3, 3, 6, 1, 0), _s(1, 1, 6, 2, 0),
    _s(2, 2, 6, 1, 4),
]
KEYS = ("ppe_nu1", "ppe_nu2", "ppe_nu_coarse", "ppe_inner", "ppe_fused", "ppe_rho_half")


def tune_ppe_schedule(solver, candidates=SCHEDULES, reps: int = 2, tol: float = None, precision: int = None,
                      margin: float = 0.02) -> dict:
    """Pick the fastest smoother schedule for `solver` (a Solver or
    BatchSolver whose pressure operator is set up, e.g. after a step) and
    apply it.  The solver's current schedule is timed first and is replaced
    only by a candidate at least `margin` faster, so timing noise does not
    flip the choice between runs.  Returns {"best": schedule, "ms":
    {"current" | index: milliseconds}}."""
    import numpy as np
    case = solver.case
    tol = case.ppe_tol if tol is None else tol
    precision = case.ppe_precision if precision is None else precision
    keep = {k: getattr(case, k) for k in KEYS}
    p_saved = solver.pull("p")
    zeros = np.zeros_like(p_saved)

    def timed(sched):
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
            return None                          # did not converge: not a candidate
        return round(best, 3)

    times = {}
    base = timed(keep)
    if base is not None:
        times["current"] = base
    for n, sched in enumerate(candidates):
        if sched == keep:
            continue
        t = timed(sched)
        if t is not None:
            times[n] = t
    solver.push("p", p_saved)
    cand = {k: v for k, v in times.items() if k != "current"}
    win = min(cand, key=cand.get) if cand else None
    if win is not None and (base is None or cand[win] < (1.0 - margin) * base):
        solver.set_params(**candidates[win])
        return {"best": dict(candidates[win]), "ms": times}
    solver.set_params(**keep)
    return {"best": keep, "ms": times}


# tiles (threads along i, along j, slab rows of the XCD order; 0: whole
# planes) tried for every stencil kernel
TILES = [(64, 4, 0), (64, 2, 0), (128, 2, 0), (256, 1, 0), (64, 4, 4), (64, 4, 8), (64, 4, 16), (128, 2, 8),
         (256, 1, 16)]


def tune_tiles(solver, candidates=TILES, steps: int = 1, margin: float = 0.02) -> dict:
    """Pick, for each tunable stencil kernel, the fastest thread-block tile
    over `steps` real steps per candidate (the steps advance the run exactly
    as untuned steps would: a tile changes no result).  A kernel keeps its
    current tile unless a candidate is at least `margin` faster.  Returns
    {"best": {region: (tx, ty, slab)}, "us": {region: {"tx x ty / slab": microseconds}}}."""
    from .types import TILE_KERNELS
    keep = {k: solver.get_tile(k) for k in TILE_KERNELS}
    order = [None] + [t for t in candidates]
    us = {r: {} for r in TILE_KERNELS.values()}
    for t in order:
        for k in TILE_KERNELS:
            solver.set_tile(k, *(keep[k] if t is None else t))
        solver.prof_enable(True)
        for _ in range(steps):
            solver.step()
        solver.sync()
        prof = solver.prof_read()
        solver.prof_enable(False)
        for k, region in TILE_KERNELS.items():
            ms, cnt = prof.get(region, (0.0, 0))
            if cnt:
                tt = keep[k] if t is None else t
                key = f"{tt[0]}x{tt[1]}/{tt[2]}"
                v = round(ms / cnt * 1e3, 3)
                us[region][key] = min(v, us[region].get(key, v))
    best = {}
    for k, region in TILE_KERNELS.items():
        cur = f"{keep[k][0]}x{keep[k][1]}/{keep[k][2]}"
        times = us[region]
        pick = keep[k]
        if times and cur in times:
            win = min(times, key=times.get)
            if times[win] < (1.0 - margin) * times[cur]:
                a, b = win.split("/")
                pick = tuple(int(x) for x in a.split("x")) + (int(b),)
        solver.set_tile(k, *pick)
        best[region] = pick
    return {"best": best, "us": us}
