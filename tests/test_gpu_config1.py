"""The 2D rising bubble of the reference's CPU configuration (64 x 128 on
[0,1]x[0,2], circle r = 0.25 at (0.5, 0.5), density ratio 1000, viscosity
ratio 100, CSF, gravity -y, no-slip walls, dt = 0.1 dx, reinitialisation
every 5 steps as the reference's interface_solver hard-codes) run for 100
steps on the device and compared with the compiled Fortran reference itself
(tests/golden/config1_ref.npz, written by tests/golden/make_config1.py with
the reference's pressure relaxation tightened to max |dp| < 1e-13).

The device path solves the pressure to ||r||/||f|| <= 1e-12, in FP64 and in
the mixed mode (FP32 V-cycles in FP64 flexible CG); the remaining
difference comes from where each pressure iteration stops.  Tolerances on
max |device - reference| / max |reference| for phi, u, v, nu, nv and p:
1e-10 after step 100 (measured <= 3.6e-11), 2e-10 after step 10 (measured
<= 1.04e-10), 2e-9 after step 1 (measured 1.05e-9: one step from rest leaves
|v| ~ 1e-3, so the reference relaxation's own stopping error, ~1e-10 of p,
is a larger share of it); the per-step mass to 1e-12 relative (measured
1.4e-15).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "config1_ref.npz")

pytestmark = pytest.mark.gpu


def _case():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_config1 import config1_case
    return config1_case()


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="tests/golden/config1_ref.npz not generated")
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_config1_100_steps_against_the_reference(precision):
    from fortrancfd_amd import Solver
    ref = np.load(FIXTURE)
    case = _case()
    case.ppe_precision = precision
    s = Solver(case)
    s.push("phi", ref["phi0"])
    s.problem_init(False)
    assert abs(s.diag()["imass"] - float(ref["imass"])) <= 1e-13 * float(ref["imass"])
    g = case.ghc
    worst = {}
    mass = []
    for n in range(1, 101):
        s.step()
        mass.append(s.diag()["mass"])
        if n in (1, 10, 100):
            for f in ("phi", "u", "v", "nu", "nv", "p"):
                a = s.pull(f)[g:-g, g:-g]
                b = ref[f"step{n}/{f}"][g:-g, g:-g]
                worst[(n, f)] = float(np.abs(a - b).max() / np.abs(b).max())
    print({f"{k[0]}/{k[1]}": f"{v:.2e}" for k, v in worst.items()})
    dm = float(np.max(np.abs(np.array(mass) - ref["mass"]) / ref["mass"]))
    print("mass history", dm)
    for step, tol in ((1, 2e-9), (10, 2e-10), (100, 1e-10)):
        assert max(v for k, v in worst.items() if k[0] == step) <= tol, (step, worst)
    assert dm <= 1e-12


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_config1_100_steps_against_the_converged_oracle(precision):
    """The same 100 steps against the oracle, whose pressure solve is direct
    (band Cholesky of the ppe_sor_init operator, exact to rounding) and which
    otherwise matches the compiled reference bit for bit.  With the device
    solving to ||r||/||f|| <= 1e-13, phi, u, v, nu, nv and p agree to 1e-10
    relative after every tenth step, from step 1 on."""
    from fortrancfd_amd import Solver
    from oracle.oracle import Oracle
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_config1 import config1_phi
    case = _case()
    case.ppe_precision = precision
    case.ppe_tol = 1e-13
    phi0 = config1_phi(case)
    Oracle.threads(1)
    s, o = Solver(case), Oracle(case)
    s.push("phi", phi0)
    o.set("phi", phi0)
    s.problem_init(False)
    o.call("problem_init", 0)
    g = case.ghc
    worst = 0.0
    for n in range(1, 101):
        s.step()
        o.call("step", 1)
        if n == 1 or n % 10 == 0:
            errs = {}
            for f in ("phi", "u", "v", "nu", "nv", "p"):
                a = s.pull(f)[g:-g, g:-g]
                b = o.get(f)[g:-g, g:-g]
                errs[f] = float(np.abs(a - b).max() / np.abs(b).max())
            print(n, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
            worst = max(worst, max(errs.values()))
            assert max(errs.values()) <= 1e-10, (n, errs)
        dm = abs(s.diag()["mass"] - o.diag()["mass"]) / o.diag()["mass"]
        assert dm <= 1e-12, (n, dm)
    print("worst", worst)
