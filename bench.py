"""Benchmark of the two-phase step on one MI355X.

Workload (default): 3D rising bubble on [0,1]^2 x [0,2] with 256^3 cells
(dx = dy = 1/256, dz = 2/256), sphere r = 0.2 at (0.5, 0.5, 0.5), density
ratio 1000, viscosity ratio 100, CSF surface tension, gravity -z, no-slip
walls, dt = 0.1 dx, reinitialisation every 5 steps, 4 Picard iterations per
step, each with a multigrid pressure solve to ||r||/||f|| <= 1e-10 (FP32
V-cycles inside FP64 defect correction).  Synthetic initial condition (signed
distance, fluid at rest); W untimed warm-up steps, then K timed steps.

Prints one JSON line: Mcell-steps/s, the roofline of the dominant kernel
(algorithmic bytes per launch / mean launch time from HIP events recorded on
the solver's stream), the FP64-vs-mixed multigrid time to 1e-10 on the last
step's pressure system, and a CPU baseline (the oracle, on a bounded sample).

`--poisson N` instead times the standalone variable-coefficient Poisson solve
on N^3 (rho = 1000 outside / 1 inside a sphere r = 0.3, smoothed over 1.5
cells; seeded U[-1,1] rhs with zero mean; Neumann walls).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s


from fortrancfd_amd.workloads import (BATCH_SIZES, batch_cases, poisson_case, rising_case,  # noqa: E402
                                     seeded_rhs, sphere_density)


def region_bytes(case, mixed: bool) -> dict:
    """Algorithmic bytes per launch of each kernel region: every distinct
    array a kernel reads or writes counted once per interior cell (ghosts,
    halo re-reads and reduction partials excluded), 8 B per FP64 value,
    4 B per FP32 value, 2 B per FP16 value (the level-0 density of the FP32
    cycle with ppe_rho_half)."""
    N = case.nx * case.ny * case.nz
    b = 4 if mixed else 8
    br = 2 if (mixed and case.ppe_rho_half) else b   # density bytes on level 0
    nchild0 = 4 if case.dz > 1.5 * case.dx else 8   # first coarsening: x,y only when dz = 2 dx
    return {
        "ls_adv_stage": 8 * N * 20 / 3,        # R{phi,nu,nv,nw,l1..} W{phi',l} averaged over 3 stages
        "redis_pass_a": 8 * N * 5,             # R{phi,S0} W{grad,s2,s3}
        "redis_pass_b": 8 * N * 8,             # R{phi,S0,grad,s2,s3,l1,l2,phi0} W{phi,l} averaged over stages
        "ccd_lines": 8 * N * 3,                # one direction per launch: R{f} W{fx,fxx}
        "curv": 8 * N * 10,
        "diff_source": 8 * N * 11,
        # R{u,v,w (Picard iterate), u,v,w old, D x3, S_old x3} W{u* x3, S x3},
        # S stored by the last of the step's picard launches only
        "predictor": 8 * N * (15 + 3 / (case.picard_iters if case.picard_iters > 0 else 4)),
        "ppe_src": 8 * N * 4,
        "projection": 8 * N * 8,
        "node_vel": 8 * N * 6,
        "mpls_update": 8 * N * 3,
        # per colour (colour-split FP32 arrays in the mixed cycle): R{e of the
        # other colour (N/2), f of this colour (N/2), rho (N, FP32 or FP16)}
        # W{e of this colour (N/2)}; averaged over the 8 launches of a V(2,2) cycle:
        # + R{ec} (N/nchild0 coarse values) in the first post-sweep half
        # (prolongation folded in) and, in the last half-sweep (CG dots
        # folded in), + R{z_old} and the residual of the pair's other cell
        # (FP32 split f, N/2, in the mixed cycle; FP64 r otherwise)
        "mg_smooth_l0": N * (8 * (1.5 * b + br) + b / nchild0 + ((0.5 * b + b) if mixed else (8 + b))) / 8,
        "mg_restrict_l0": N * (2 * b + br + b * 2.0 / nchild0),   # R{e,f,rho} W{fc,ec}
        "mg_resid_l0": N * (3 * 8 + (2 * 4 if mixed else 0)),  # R{p,f,rho} (W{r32,e32})
        # flexible CG, averaged over its two launches per iteration:
        # dir R{z, p, q, x, rho} W{p, q, x, z_old}; update R{q, r} W{r, r32}
        "fcg_vectors": N * ((8 * 7 + 2 * b) + (8 * 3 + (4 if mixed else 0))) / 2,
    }


FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (every lane an FMA each cycle)


def region_flops(case) -> dict:
    """Algorithmic FP64 flops per launch of the FP64-ALU-bound regions (a
    divide, square root or sine counts as one flop).  One WENO5-JS
    reconstruction as the kernels write it: three smoothness indicators
    (~33), q_k = (EPS + beta_k)^2 (6), the product-form weights (5), their
    normalisation (3 + 3) and the stencil combination (~25): ~75.
    WENO stage: 3 faces per cell (each shared by two cells) x 2
    reconstructions = 450, flux splits 3 x 7, face sums, divergence and the
    increment-form update ~20: ~490 per cell.  Reinit pass A: 2 one-sided
    reconstructions per direction = 450, 7 backward differences per
    direction x 2 = 42, Godunov selection and |grad| ~15, smoothed
    Heaviside / delta and the lambda terms ~20: ~530 per cell."""
    N = case.nx * case.ny * case.nz
    return {"ls_adv_stage": N * 490, "redis_pass_a": N * 530}


def profiled_traffic(region: str):
    """HBM bytes per launch of a kernel region as measured by the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/*.txt, line
    `traffic <region> ... bytes_per_launch=B`, counters scaled by a
    device-to-device copy of known size); None when no profile holds it."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".txt"):
            continue
        with open(os.path.join(pdir, fn)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) > 2 and parts[0] == "traffic" and parts[1] == region:
                    kv = dict(p.split("=", 1) for p in parts[2:] if "=" in p)
                    return float(kv["bytes_per_launch"]), fn
    return None, None


def cpu_baseline_reference(n: int = 40, steps: int = 3, threads: int = 0) -> dict:
    """The compiled Fortran reference (oracle/_ref/ref3d, its own OpenMP
    build) on a bounded sample: the rising bubble on a uniform n x n x 2n grid
    of [0,1]^2 x [0,2] (the reference reads cells per unit length, so the
    dz = 2 dx grid of the GPU workload has no reference input), problem_init
    untimed, then `steps` main-loop steps in one process, timed inside it
    around the step loop only (no field I/O), with the reference's own
    pressure relaxation (max |dp| < 1e-7, p_w1 = 1.5).  Threads: the box's
    CPU share for one GPU (16), bound to cores."""
    from oracle import refrun
    from fortrancfd_amd.types import Case, bc_copy, bubble_phi
    if not refrun.available(3):
        raise RuntimeError("reference harness not built")
    th = threads or min(16, os.cpu_count() or 1)
    case = Case(dim=3, nx=n, ny=n, nz=2 * n, xend=1.0, yend=1.0, zend=2.0, gvec=(0.0, 0.0, -1.0),
                cfl=0.1, cfl_r=0.1, redis_interval=5, picard_iters=4, name="rising_bubble_3d")
    phi = bc_copy(case, bubble_phi(case, 0.2, (0.5, 0.5, 0.5)))
    out = refrun.run(case, "init_nore", {"phi": phi}, threads=th)
    state = {k: out[k] for k in refrun.HARNESS_FIELDS_3D}
    r = refrun.run_steps_timed(case, state, steps, out["_imass"], th)
    dt = r["seconds"]
    return {"value": round(case.ncells * steps / dt / 1e6, 6), "unit": "Mcell-steps/s", "cores": th,
            "kind": "reference",
            "sample": f"compiled Fortran reference (newOpenMP 3D, flang -O2 -fopenmp) main loop on the rising bubble "
                      f"at {n}x{n}x{2 * n} (uniform grid), {steps} steps timed inside the program (no file I/O), "
                      f"its own pressure relaxation to max|dp| < 1e-7 (p%glb%piter {r['piter']} at the end), {th} OpenMP "
                      f"threads bound to cores, {dt:.1f} s"}


def cpu_baseline(steps: int = 5, n: int = 96, threads: int = 0) -> dict:
    """The oracle (CPU restatement of the reference step, OpenMP) on a bounded
    sample: the same rising-bubble case on n^3 cells, init + `steps` steps.
    Its pressure solve is a converged Jacobi-PCG to the same 1e-10 relative
    residual, because the reference's relaxation cannot reach 1e-10."""
    from oracle.oracle import Oracle
    from fortrancfd_amd.types import bc_copy, bubble_phi
    th = threads or min(16, os.cpu_count() or 1)
    used = Oracle.threads(th)
    case = rising_case(n, 0, 1e-10)
    o = Oracle(case)
    o.set("phi", bc_copy(case, bubble_phi(case, 0.2, (0.5, 0.5, 0.5))))
    o.call("problem_init", 0)
    o.call("step", 1)   # first step from p = 0: untimed
    t = time.perf_counter()
    for _ in range(steps):
        o.call("step", 1)
    dt = time.perf_counter() - t
    return {"value": round(case.ncells * steps / dt / 1e6, 4), "unit": "Mcell-steps/s", "cores": used,
            "kind": "port",
            "sample": f"oracle step (reference operation order, Jacobi-PCG pressure solve to 1e-10) on the "
                      f"same rising-bubble case at {n}x{n}x{n}, {steps} timed steps after 1 untimed step, "
                      f"{used} OpenMP threads, {dt:.1f} s"}


MG_VARIANTS = [  # (name, precision, krylov, V-cycles per outer iteration, pre/post sweeps[, FP16 density])
    ("fp64_plain", 0, 0, 1, 2, 2), ("fp64_fcg", 0, 1, 1, 2, 2), ("fp64_fcg2", 0, 1, 2, 2, 2),
    ("mixed_plain", 1, 0, 1, 2, 2), ("mixed_plain2", 1, 0, 2, 2, 2), ("mixed_fcg", 1, 1, 1, 2, 2),
    ("mixed_fcg2", 1, 1, 2, 2, 2),
    # the FP32 cycle with every array in FP32 (no FP16 density / face coefficients)
    ("mixed_fcg_fp32", 1, 1, 1, 2, 2, 0), ("mixed_plain_fp32", 1, 0, 1, 2, 2, 0),
]


def mg_compare(s, case, variants=MG_VARIANTS) -> dict:
    """FP64 vs mixed multigrid to 1e-10 on one pressure system (the operator
    of rho now and the rhs in PPE_SRC), every variant from p = 0; the best
    FP64 and the best mixed variant are compared.  Restores the case's own
    solver settings afterwards."""
    keep = (case.ppe_krylov, case.ppe_inner, case.ppe_nu1, case.ppe_nu2, case.ppe_rho_half)
    out = {"variants": {}}
    zeros = np.zeros(case.size)
    sols = {}
    for name, prec, kry, inner, nu1, nu2, *half in variants:
        s.set_params(ppe_krylov=kry, ppe_inner=inner, ppe_nu1=nu1, ppe_nu2=nu2,
                     ppe_rho_half=half[0] if half else keep[4])
        best = None
        for rep in range(2):
            s.push("p", zeros)
            s.sync()
            t = time.perf_counter()
            cyc, rel = s.ppe_solve_rhs(1e-10, prec)
            s.sync()
            ms = (time.perf_counter() - t) * 1e3
            best = ms if best is None else min(best, ms)
        sols[name] = case.interior(s.pull("p")).copy()
        out["variants"][name] = {"ms": round(best, 3), "outer_iters": cyc, "rel_resid": rel}
    s.set_params(ppe_krylov=keep[0], ppe_inner=keep[1], ppe_nu1=keep[2], ppe_nu2=keep[3], ppe_rho_half=keep[4])
    conv = {k: v for k, v in out["variants"].items() if v["rel_resid"] <= 1e-10}
    f64 = min((k for k in conv if k.startswith("fp64")), key=lambda k: conv[k]["ms"], default=None)
    mix = min((k for k in conv if k.startswith("mixed")), key=lambda k: conv[k]["ms"], default=None)
    m32 = min((k for k in conv if k.startswith("mixed") and k.endswith("_fp32")), key=lambda k: conv[k]["ms"],
              default=None)
    if f64 and m32:
        # every array of the FP32 cycle in FP32 (the precision the mixed scheme names)
        out["speedup_fp32_only"] = round(conv[f64]["ms"] / conv[m32]["ms"], 3)
        out["mixed_fp32_only_best"] = m32
    if f64 and mix:
        a, b = sols[f64], sols[mix]
        out.update({"fp64_best": f64, "fp64_ms": conv[f64]["ms"], "mixed_best": mix, "mixed_ms": conv[mix]["ms"],
                    "speedup": round(conv[f64]["ms"] / conv[mix]["ms"], 3),
                    "p_rel_diff": float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))})
    if "fp64_plain" in conv and "mixed_plain" in conv:
        # the same V(2,2) cycle iterated in both precisions (no Krylov acceleration)
        out["speedup_plain_vcycle"] = round(conv["fp64_plain"]["ms"] / conv["mixed_plain"]["ms"], 3)
    return out


def config2_run(n: int, steps: int = 200) -> dict:
    """The 3D rising bubble of the benchmark from rest for `steps` steps (the
    whole configuration), once with the FP64 multigrid and once with FP32
    V-cycles inside the FP64 outer iteration: wall time, pressure iterations
    per step, the mixed run's final u, v, w, p, phi against the FP64 run's
    (the FP64 path is pinned to the oracle at 1e-10 by the tests), and the
    bubble volume drift sum(1 - H) dV of each run."""
    from fortrancfd_amd import Solver
    from fortrancfd_amd.types import PPE_FP64, PPE_MIXED, bc_copy, bubble_phi
    runs, fields = {}, {}
    for name, prec in (("fp64", PPE_FP64), ("mixed", PPE_MIXED)):
        case = rising_case(n, prec, 1e-10)
        s = Solver(case)
        s.push("phi", bc_copy(case, bubble_phi(case, 0.2, (0.5, 0.5, 0.5))))
        s.problem_init(False)
        d0 = s.diag()
        vol_tot = float(case.nx * case.ny * case.nz) * case.dx * case.dy * case.dz
        bub0 = vol_tot - d0["vol"]
        cyc = []
        s.sync()
        t = time.perf_counter()
        for _ in range(steps):
            s.step()
            cyc.append(s.diag()["ppe_cycles_step"])
        s.sync()
        dt = time.perf_counter() - t
        d = s.diag()
        bub = vol_tot - d["vol"]
        runs[name] = {"seconds": round(dt, 3), "ms_per_step": round(dt * 1e3 / steps, 3),
                      "ppe_iters_per_step": {"min": int(min(cyc)), "mean": round(float(np.mean(cyc)), 2),
                                             "max": int(max(cyc))},
                      "bubble_volume_drift": float(bub / bub0 - 1.0), "mass_drift": float(d["mass"] / d0["mass"] - 1.0)}
        fields[name] = {f: case.interior(s.pull(f)).copy() for f in ("u", "v", "w", "p", "phi")}
        s.close()
        print(f"config-2 {name}: {steps} steps in {dt:.2f} s", file=sys.stderr, flush=True)
    err = {f: float(np.abs(fields["mixed"][f] - fields["fp64"][f]).max() / np.abs(fields["fp64"][f]).max())
           for f in fields["fp64"]}
    return {"steps": steps, "grid": [n, n, n], "runs": runs, "mixed_vs_fp64_max_rel": err}


def poisson_bench(n: int) -> dict:
    from fortrancfd_amd import Solver
    case = poisson_case(n, ppe_rho_half=1)
    s = Solver(case, allow_unconverged=True)
    s.push("rho", sphere_density(case))
    s.push("ppe_src", seeded_rhs(case))
    s.ppe_sor_init()
    res = mg_compare(s, case)
    return {"metric": f"MG ms to 1e-10, {n}^3 variable-coefficient Poisson", "value": res.get("mixed_ms"),
            "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f64 (FP32 V-cycles in FP64 "
            "defect correction)", "data": "synthetic", "config": {"workload": f"poisson_{n}^3"}, "mg": res}


def adjoint_bench(n: int, nsteps: int = 300, every: int = 20, cfl: float = 0.3) -> dict:
    """The adjoint advection workload on n^3 cells of [0,1]^3: deformation
    velocity of amplitude A = 1 and period T = 3, sphere r = 0.15 at
    (0.35, 0.35, 0.35), nsteps WENO5 SSP-RK3 steps of dt = cfl dx in FP64,
    J = L2 norm of phi_T - phi_target with phi_target from A = 1.05.
    Times the forward (checkpoint every `every` steps) and the backward
    (recompute + hand-written adjoint), and checks dJ/dA and dJ/dphi0 along a
    smooth direction against central finite differences with eps = 1e-6;
    then times the same forward in FP32 and reports its deviation from the
    FP64 phi_T and J."""
    import torch
    from fortrancfd_amd import Solver
    from fortrancfd_amd.adjoint import advect, advect_f32, l2_loss
    from fortrancfd_amd.types import Case, bc_copy, bubble_phi
    case = Case(dim=3, nx=n, ny=n, nz=n, xend=1.0, yend=1.0, zend=1.0, cfl=cfl, name="adjoint")
    s = Solver(case)
    dev = torch.device("cuda")
    T = 3.0
    phi0 = torch.from_numpy(bc_copy(case, bubble_phi(case, 0.15, (0.35, 0.35, 0.35))).reshape(-1)).to(dev)

    def fwd(p, a):
        with torch.no_grad():
            return advect(p, torch.tensor(a, dtype=torch.float64, device=dev), s, nsteps, T, nsteps)

    target = fwd(phi0, 1.05)
    p0 = phi0.clone().requires_grad_(True)
    A = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phiT = advect(p0, A, s, nsteps, T, every)
    J = l2_loss(phiT, target, case)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    eps = 1e-6
    # smooth direction for phi0: a low Fourier mode over the box
    X, Y, Z = case.coords()
    v = torch.from_numpy(np.ascontiguousarray(np.sin(np.pi * X) * np.sin(2 * np.pi * Y) * np.cos(np.pi * Z)).reshape(-1)).to(dev)
    del X, Y, Z
    loss = lambda p, a: float(l2_loss(fwd(p, a), target, case))
    fd_A = (loss(phi0, 1.0 + eps) - loss(phi0, 1.0 - eps)) / (2 * eps)
    fd_v = (loss(phi0 + eps * v, 1.0) - loss(phi0 - eps * v, 1.0)) / (2 * eps)
    ad_v = float((p0.grad * v).sum())
    ad_A = float(A.grad)
    cells = case.ncells
    # FP32 forward (one untimed step allocates its scratch)
    advect_f32(phi0, 1.0, s, 1, T)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    p32 = advect_f32(phi0, 1.0, s, nsteps, T)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    g = case.ghc
    inner = lambda a: a.detach().reshape(case.shape)[g:-g, g:-g, g:-g].double()
    err32 = float((inner(p32) - inner(phiT)).abs().max() / inner(phiT).abs().max())
    J32 = float(l2_loss(p32.double(), target, case))
    return {"metric": f"adjoint level-set advection {n}^3, {nsteps} steps: gradient vs central FD (eps 1e-6)",
            "value": max(abs(ad_A - fd_A) / abs(fd_A), abs(ad_v - fd_v) / abs(fd_v)), "unit": "max rel error",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"adjoint_deformation_{n}^3", "steps": nsteps, "checkpoint_every": every,
                       "dt_over_dx": cfl, "period": T, "A": 1.0, "A_target": 1.05},
            "loss": float(J.detach()), "dJdA": ad_A, "fd_dJdA": fd_A, "dJdphi_dir": ad_v, "fd_dJdphi_dir": fd_v,
            "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
            "forward_Mcell_steps_per_s": round(cells * nsteps / (t1 - t0) / 1e6, 1),
            "backward_Mcell_steps_per_s": round(cells * nsteps / (t2 - t1) / 1e6, 1),
            "fp32_forward_s": round(t4 - t3, 3),
            "fp32_forward_Mcell_steps_per_s": round(cells * nsteps / (t4 - t3) / 1e6, 1),
            "fp32_phiT_max_rel_dev": err32, "fp32_loss": J32,
            "fp32_loss_rel_dev": abs(J32 - float(J.detach())) / float(J.detach())}


def batch_bench(ncases: int = 64, steps: int = 500, warmup: int = 2, loop_steps: int = 5,
                precision: int = 1, serial: bool = False, regions: bool = False) -> dict:
    """Aggregate throughput of the batched 2D workload: one batched context
    per grid size, each stepped by its own host thread (its own HIP stream),
    against the same cases stepped one after another in ordinary contexts
    (a bounded sample of loop_steps steps per case).  serial: the contexts
    are stepped one after another from the main thread instead.  regions:
    afterwards each context steps 5 more times with the library's HIP-event
    regions on (graphs off) and their ms per step are reported per grid."""
    import threading
    from fortrancfd_amd import Solver
    from fortrancfd_amd.solver import BatchSolver
    from fortrancfd_amd.types import bc_copy, bubble_phi
    cases = batch_cases(ncases, precision=precision)
    groups = {}
    for c, r in cases:
        groups.setdefault((c.nx, c.ny), []).append((c, r))
    solvers = []
    for key in sorted(groups):
        g = groups[key]
        b = BatchSolver([c for c, _ in g])
        b.push("phi", np.stack([bc_copy(c, bubble_phi(c, r, (0.5, 0.5))) for c, r in g]))
        b.problem_init(False)
        solvers.append(b)
    cells = sum(c.nx * c.ny for c, _ in cases)

    def run(nsteps):
        errs = []

        def worker(b):
            try:
                for _ in range(nsteps):
                    b.step()
                b.sync()
            except Exception as e:   # surfaced below, with the cases that are no longer finite
                bad = [q for q, a in enumerate(b.pull("phi")) if not np.isfinite(a).all()]
                bad += [q for q, a in enumerate(b.pull("p")) if not np.isfinite(a).all()]
                errs.append(RuntimeError(f"{e}; grid {b.case.nx}x{b.case.ny}, non-finite cases "
                                         f"{sorted(set(bad))}: " + ", ".join(
                                             f"rho_12={b.cases[q].rho_12:.4g} we={b.cases[q].we:.3g}"
                                             for q in sorted(set(bad)))))
        t = time.perf_counter()
        if serial:
            for b in solvers:
                worker(b)
        else:
            th = [threading.Thread(target=worker, args=(b,)) for b in solvers]
            for x in th:
                x.start()
            for x in th:
                x.join()
        dt = time.perf_counter() - t
        if errs:
            raise errs[0]
        return dt

    run(warmup)
    # smoother schedule per grid, tuned on each context's current pressure system
    from fortrancfd_amd.autotune import tune_ppe_schedule
    tuned = {}
    for b in solvers:
        res = tune_ppe_schedule(b)
        tuned[f"{b.case.nx}x{b.case.ny}"] = {"schedule": res["best"], "candidate_ms": res["ms"]}
    run(1)   # captures the chosen schedule's cycle graphs
    t_batch = run(steps)
    mass_drift = max(float(np.max(np.abs(d["mass"] - d["imass"]) / d["imass"]))
                     for d in (b.case_diag() for b in solvers))
    # per-case launch loop on a bounded sample
    singles = []
    for c, r in cases:
        s = Solver(c)
        s.push("phi", bc_copy(c, bubble_phi(c, r, (0.5, 0.5))))
        s.problem_init(False)
        for _ in range(warmup):
            s.step()
        singles.append(s)
    for s in singles:
        s.sync()
    t = time.perf_counter()
    for s in singles:
        for _ in range(loop_steps):
            s.step()
        s.sync()
    t_loop = time.perf_counter() - t
    rate_b = cells * steps / t_batch / 1e6
    rate_l = cells * loop_steps / t_loop / 1e6
    region_ms = {}
    if regions:
        for b in solvers:
            b.prof_enable(True)
            for _ in range(5):
                b.step()
            b.sync()
            pr = b.prof_read()
            b.prof_enable(False)
            region_ms[f"{b.case.nx}x{b.case.ny}"] = {k: [round(v[0] / 5, 4), v[1] // 5] for k, v in pr.items() if v[1]}
    return {"metric": "batched 2D rising bubbles: aggregate Mcell-steps/s", "value": round(rate_b, 2),
            "unit": "Mcell-steps/s", "higher_is_better": True, "n_gpus": 1,
            "dtype": "f64 (FP32 V-cycles in FP64 flexible CG)" if precision else "f64", "data": "synthetic",
            "config": {"workload": f"batched_2d_rising_bubbles_{ncases}", "cases": ncases, "steps": steps,
                       "host_threads": 1 if serial else len(solvers),
                       "grids": {f"{k[0]}x{k[1]}": len(v) for k, v in sorted(groups.items())},
                       "dt_over_dx": {f"{k[0]}x{k[1]}": round(v[0][0].cfl, 5) for k, v in sorted(groups.items())},
                       "total_cells": cells},
            "ms_per_step": round(t_batch / steps * 1e3, 3),
            "per_case_loop": {"value": round(rate_l, 2), "unit": "Mcell-steps/s", "steps_per_case": loop_steps,
                              "seconds": round(t_loop, 3)},
            "speedup_vs_per_case_loop": round(rate_b / rate_l, 2), "autotuned": tuned,
            **({"region_ms_per_step_and_launches": region_ms} if regions else {}),
            "max_rel_mass_drift": mass_drift}


def m3_bench(unit_grid: int, steps: int, warmup: int, precision: int = 1) -> dict:
    """The MPLS_RisingBubble two-bubble step at unit_grid cells per unit
    length (35 is the reference's 140 x 140 x 280): setup (reinitialisation,
    first momentum source, mass targets) untimed, warmup steps, then `steps`
    main-loop steps timed, then the same steps again with the event regions."""
    from fortrancfd_amd.solver import M3Solver
    from fortrancfd_amd.workloads import m3_bubbles, m3_case
    case = m3_case(unit_grid, precision)
    s = M3Solver(case)
    for k, a in m3_bubbles(case).items():
        s.push(k, a)
    s.set_time(0.0, 0)
    t0 = time.perf_counter()
    s.setup()
    s.sync()
    t_setup = time.perf_counter() - t0
    st0 = s.state()
    for w in range(warmup):
        s.step()
        print(f"m3 warmup step {w + 1}/{warmup} picard {s.state()['picard']}", file=sys.stderr, flush=True)
    s.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    s.sync()
    t1 = time.perf_counter()
    st = s.state()
    picard, cycles = [], []
    s.prof_enable(True)
    tp0 = time.perf_counter()
    for _ in range(steps):
        s.step()
        picard.append(s.state()["picard"])
        cycles.append(s.diag()["ppe_cycles_step"])
    s.sync()
    tp1 = time.perf_counter()
    prof = s.prof_read()
    s.prof_enable(False)
    regions = {k: {"ms_per_step": round(ms / steps, 3), "launches_per_step": round(cnt / steps, 1)}
               for k, (ms, cnt) in prof.items() if cnt}
    s.close()
    value = case.ncells * steps / (t1 - t0) / 1e6
    return {
        "metric": "Mcell-steps/s, MPLS_RisingBubble two-bubble step", "value": round(value, 3),
        "unit": "Mcell-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / steps, 3), "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (the reference's initial condition)",
        "config": {"workload": f"mpls_rising_bubble_{case.nx}x{case.ny}x{case.nz}",
                   "grid": [case.nx, case.ny, case.nz], "domain": [4.0, 4.0, 8.0],
                   "ppe": ("fp32 V-cycles in fp64 flexible CG" if precision else "fp64") + " to 1e-10"},
        "setup_s": round(t_setup, 3),
        "picard_per_step_profiled_run": picard, "ppe_cycles_per_step_profiled_run": cycles,
        "mass_drift": {"total": abs(st["mass"] - st0["imass"]) / st0["imass"]},
        "bubble_imass": list(st0["bubble_imass"]),
        "profiled_ms_per_step": round((tp1 - tp0) * 1e3 / steps, 3),
        "regions": regions,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--poisson", type=int, default=0)
    ap.add_argument("--adjoint", type=int, default=0, help="run the adjoint advection workload on N^3 cells")
    ap.add_argument("--adjoint-steps", type=int, default=300)
    ap.add_argument("--batch", type=int, default=0, help="run the batched 2D workload with N cases")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-serial", action="store_true", help="step the batched contexts from one host thread")
    ap.add_argument("--batch-regions", action="store_true", help="add the per-grid HIP-event region breakdown")
    ap.add_argument("--regions-out", default="")
    ap.add_argument("--no-tune", action="store_true", help="keep the default smoother schedule")
    ap.add_argument("--no-config2", action="store_true", help="skip the 200-step FP64 vs mixed run")
    ap.add_argument("--config2-steps", type=int, default=200)
    ap.add_argument("--m3", type=int, default=0, help="run the MPLS_RisingBubble step at N cells per unit length")
    ap.add_argument("--krylov", type=int, default=1)
    ap.add_argument("--inner", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU decomposition "
              "is a design note, not a measured path.", file=sys.stderr)
        return 2
    if args.poisson:
        print(json.dumps(poisson_bench(args.poisson)))
        return 0
    if args.adjoint:
        print(json.dumps(adjoint_bench(args.adjoint, args.adjoint_steps)))
        return 0
    if args.m3:
        print(json.dumps(m3_bench(args.m3, args.steps, args.warmup, 1 if args.precision == "mixed" else 0)))
        return 0
    if args.batch:
        print(json.dumps(batch_bench(args.batch, args.batch_steps, precision=1 if args.precision == "mixed" else 0,
                                     serial=args.batch_serial, regions=args.batch_regions)))
        return 0

    from fortrancfd_amd import Solver
    from fortrancfd_amd.types import PPE_FP64, PPE_MIXED, bc_copy, bubble_phi
    mixed = args.precision == "mixed"
    case = rising_case(args.n, PPE_MIXED if mixed else PPE_FP64, 1e-10)
    case.ppe_krylov, case.ppe_inner = args.krylov, args.inner
    s = Solver(case)   # an unconverged pressure solve raises
    s.push("phi", bc_copy(case, bubble_phi(case, 0.2, (0.5, 0.5, 0.5))))
    s.problem_init(False)
    for w in range(args.warmup):
        s.step()
        print(f"warmup step {w + 1}/{args.warmup} {s.diag()}", file=sys.stderr, flush=True)
    tuned = tiles = None
    if not args.no_tune:
        # thread-block tiles of the stencil kernels (untimed steps that
        # advance the run as any step does), then the smoother schedule for
        # this system (untimed; the pressure is restored)
        from fortrancfd_amd.autotune import tune_ppe_schedule, tune_tiles
        tiles = tune_tiles(s)
        print(f"autotuned tiles {tiles['best']}", file=sys.stderr, flush=True)
        tuned = tune_ppe_schedule(s)
        print(f"autotuned schedule {tuned['best']} ({tuned['ms']} ms)", file=sys.stderr, flush=True)
        s.step()   # captures the chosen schedule's cycle graphs
    s.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step()
    s.sync()
    t1 = time.perf_counter()
    d = s.diag()
    # the pressure iterations of each step of the profiled repeat below (a
    # diag read synchronises, so it stays out of the timed region)
    cyc_prof = []
    print(f"timed {args.steps} steps in {t1 - t0:.3f} s; last diag {d}", file=sys.stderr, flush=True)
    # the same number of steps again with a HIP event pair around every kernel
    # launch on the solver's stream (V-cycles then launch directly instead of
    # replaying their graph, so this repeat is a little slower end to end)
    s.prof_enable(True)
    tp0 = time.perf_counter()
    for k in range(args.steps):
        s.step()
        cyc_prof.append(s.diag()["ppe_cycles_step"])
    s.sync()
    tp1 = time.perf_counter()
    prof = s.prof_read()
    s.prof_enable(False)
    print(f"profiled {args.steps} steps in {tp1 - tp0:.3f} s", file=sys.stderr, flush=True)
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = case.ncells * args.steps / (t1 - t0) / 1e6

    rb = region_bytes(case, mixed)
    regions = {}
    tot_ms, tot_bytes = 0.0, 0.0
    for name, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        regions[name] = {"ms_total": round(ms, 3), "launches": cnt, "ms_per_launch": round(ms / cnt, 5)}
        tot_ms += ms
        if name in rb:
            gbs = rb[name] / (ms / cnt * 1e-3) / 1e9
            regions[name]["GB/s"] = round(gbs, 1)
            tot_bytes += rb[name] * cnt
    # the roofline is quoted for the costliest region whose algorithmic bytes
    # are defined: the small coarse multigrid levels (mg_smooth_coarse etc.)
    # are launch-bound and L2-resident, and the event pairs around their
    # ~4 us launches inflate their profiled share
    dom_any = max(regions, key=lambda k: regions[k]["ms_total"])
    dom = max((k for k in regions if k in rb), key=lambda k: regions[k]["ms_total"])
    dom_r = regions[dom]
    achieved = dom_r.get("GB/s")
    traffic, tsrc = profiled_traffic(dom) if case.nx == 256 else (None, None)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": round(traffic) if traffic else None,
                "bytes_per_launch": rb.get(dom), "ms_per_launch": dom_r["ms_per_launch"],
                "share_of_step": round(dom_r["ms_total"] / ((tp1 - tp0) * 1e3), 4)}
    if tsrc:
        roofline["traffic_source"] = "profiles/" + tsrc
    if dom_any != dom:
        roofline["costliest_region_profiled"] = dom_any
    # the FP64-ALU-bound regions against the FP64 vector peak (their launch
    # means include the reinit launches queued after its stop, a few per run)
    alu = {}
    for name, fl in region_flops(case).items():
        if name in regions:
            tf = fl / (regions[name]["ms_per_launch"] * 1e-3) / 1e12
            alu[name] = {"bound": "fp64-valu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tf / FP64_PEAK_TFLOPS, 4), "flops_per_launch": fl,
                         "ms_per_launch": regions[name]["ms_per_launch"]}
    out = {
        "metric": "Mcell-steps/s @256^3 two-phase, %HBM peak; MG ms to 1e-10 mixed vs FP64",
        "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"rising_bubble_3d_{case.nx}x{case.ny}x{case.nz}", "grid": [case.nx, case.ny, case.nz],
                   "domain": [1.0, 1.0, 2.0], "ppe": ("fp32" if mixed else "fp64") + f" V({case.ppe_nu1},{case.ppe_nu2}) x{case.ppe_inner}" + (" preconditioning fp64 flexible CG" if args.krylov else (" in fp64 defect correction" if mixed else "")),
                   "ppe_tol": 1e-10, "picard": 4, "redis_interval": 5},
        "roofline": roofline,
        "alu_roofline": alu,
        "step_hbm_frac_measured_kernels": round(tot_bytes / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms_per_step": round(tot_ms / args.steps, 3),
        "profiled_ms_per_step": round((tp1 - tp0) * 1e3 / args.steps, 3),
        "ppe_cycles_last_step": d["ppe_cycles_step"],
        "ppe_schedule": {k: getattr(case, k) for k in ("ppe_nu1", "ppe_nu2", "ppe_nu_coarse", "ppe_inner",
                                                        "ppe_fused", "ppe_rho_half")},
    }
    if tuned:
        out["autotune_candidate_ms"] = tuned["ms"]
    if tiles:
        out["tiles"] = {k: f"{v[0]}x{v[1]}/{v[2]}" for k, v in tiles["best"].items()}
        out["tile_candidate_us"] = tiles["us"]
    if not args.no_mg:
        out["mg"] = mg_compare(s, case)
    out["ppe_iters_per_step_profiled_run"] = {"min": int(min(cyc_prof)), "mean": round(float(np.mean(cyc_prof)), 2),
                                              "max": int(max(cyc_prof))}
    if args.regions_out:
        with open(args.regions_out, "w") as fh:
            json.dump(regions, fh, indent=1)
    s.close()
    if not args.no_config2 and case.nx == 256:
        out["config2"] = config2_run(case.nx, args.config2_steps)
    if not args.no_cpu:
        # the reference's own OpenMP build; the C restatement (same operation
        # order, pressure solved to the same 1e-10) is reported beside it
        try:
            out["cpu_baseline"] = cpu_baseline_reference()
        except Exception as e:  # the reference build is optional at run time
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        try:
            out["cpu_port"] = cpu_baseline()
        except Exception as e:
            out["cpu_port"] = {"value": None, "error": str(e)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
