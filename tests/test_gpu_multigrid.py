"""Multigrid paths that only grids of realistic size reach: the fused
multi-sweep smoother of the small coarse levels (two ping-ponged error
buffers) and the LDS/register top level of the single-workgroup tail.

At 128 x 128 x 256 cells (dz = 2 dx) the hierarchy is 128x128x256,
64x64x256, 32x32x128 (fused), 16x16x64 (fused), 8x8x32 (tail top),
4x4x16 (tail), 2x2x8 (dense).  The FP32 cycle runs through all of them; the
FP64 cycle uses single-colour launches on every level and no tail top.
Both preconditioners describe the same V(2,2) cycle, so the flexible-CG
iteration counts must agree (+1 for FP32 round-off) and the two pressure
fields must agree to the solve tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n=128):
    from fortrancfd_amd import Solver
    from fortrancfd_amd.types import PPE_MIXED, Case, bc_copy, bubble_phi
    case = Case(dim=3, nx=n, ny=n, nz=2 * n, xend=1.0, yend=1.0, zend=2.0, gvec=(0.0, 0.0, -1.0),
                cfl=0.1, cfl_r=0.1, ppe_tol=1e-10, ppe_precision=PPE_MIXED, redis_interval=5,
                picard_iters=2, ppe_max_cycles=100, name="rising_bubble_3d")
    s = Solver(case, allow_unconverged=True)
    s.push("phi", bc_copy(case, bubble_phi(case, 0.2, (0.5, 0.5, 0.5))))
    s.problem_init(False)
    s.step()
    return case, s


def test_fused_coarse_smoothing_matches_fp64_cycle():
    from fortrancfd_amd.types import PPE_FP64, PPE_MIXED
    case, s = _system()
    zeros = np.zeros(case.size)
    out = {}
    for name, prec in (("fp64", PPE_FP64), ("mixed", PPE_MIXED)):
        s.set_params(ppe_krylov=1, ppe_inner=1)
        for rep in range(2):   # the second solve replays the captured graphs
            s.push("p", zeros)
            cyc, rel = s.ppe_solve_rhs(1e-10, prec)
            print(name, rep, cyc, rel)
            assert rel <= 1e-10, (name, rep, rel)
            assert cyc <= 20, (name, rep, cyc)
        out[name] = (cyc, case.interior(s.pull("p")).copy())
    assert out["mixed"][0] <= out["fp64"][0] + 1
    a, b = out["fp64"][1], out["mixed"][1]
    assert np.abs(a - b).max() <= 1e-8 * np.abs(a).max()


@pytest.mark.parametrize("fused", [0, 4])
def test_fused_coarse_smoothing_plain_cycles(fused):
    """Defect correction (no Krylov): every cycle's output is used as is, so a
    wrong coarse-level sweep shows as a stalled or slower iteration.  fused =
    4 also runs the second level (64x64x256) through the fused tiles."""
    from fortrancfd_amd.types import PPE_FP64, PPE_MIXED
    case, s = _system()
    zeros = np.zeros(case.size)
    cyc = {}
    for name, prec in (("fp64", PPE_FP64), ("mixed", PPE_MIXED)):
        s.set_params(ppe_krylov=0, ppe_inner=1, ppe_fused=fused)
        s.push("p", zeros)
        c, rel = s.ppe_solve_rhs(1e-10, prec)
        print(name, c, rel)
        assert rel <= 1e-10, (name, rel)
        cyc[name] = c
    assert cyc["mixed"] <= cyc["fp64"] + 1


def test_every_tuner_schedule_reaches_the_same_pressure():
    """Each smoother schedule the autotuner may choose (sweep counts, coarse
    FP32 tiling 0..3, FP16 density in the FP32 level-0 sweeps, two V-cycles
    per outer iteration) is a preconditioner of the same FP64 flexible CG:
    every one must reach the tolerance and the same pressure to it."""
    from fortrancfd_amd.autotune import SCHEDULES
    from fortrancfd_amd.types import PPE_MIXED
    case, s = _system()
    zeros = np.zeros(case.size)
    ref = None
    g = case.ghc
    for sched in SCHEDULES:
        s.set_params(**sched)
        s.push("p", zeros)
        cyc, rel = s.ppe_solve_rhs(1e-10, PPE_MIXED)
        p = s.pull("p")[g:-g, g:-g, g:-g]
        print(sched, cyc, rel)
        assert rel <= 1e-10 and cyc <= 20, (sched, cyc, rel)
        if ref is None:
            ref = p
        else:
            d = np.abs(p - ref).max() / np.abs(ref).max()
            assert d <= 1e-8, (sched, d)
