"""The drop-in, end to end: the reference's own newOpenMP 2D data modules and
main loop (`Two phase flows/newOpenMP/2D/src/main.f90:13-46`) with switch,
interface_solver, ns_solver and ls_mv answered by libfcfd through the fcfd_c
binding module of the integration notes (oracle/n2_dropin.F90, built by
oracle/build_ref.sh into oracle/_ref/n2_dropin).

The program reads the reference-format input file itself (its manager_init
derives dt, rdt and ls_wid) and the 2D rising bubble of
tests/golden/config1_ref.npz (64 x 128, density ratio 1000, 100 steps, the
reference's pressure relaxation tightened to max |dp| < 1e-13).  Tolerances
are those of the Python-side config-1 test: max |drop-in - reference| /
max |reference| <= 1e-10 for phi, u, v, p, nu, nv after step 100 and the
per-step mass (written by the program from its ls_mv shim) to 1e-12 relative.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "config1_ref.npz")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "n2_dropin")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="tests/golden/config1_ref.npz not generated")
@pytest.mark.skipif(not os.access(DROPIN, os.X_OK), reason="oracle/_ref/n2_dropin not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_dropin_main_loop_matches_the_reference(precision):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_config1 import config1_case
    ref = np.load(FIXTURE)
    case = config1_case()
    nsteps = 100
    with tempfile.TemporaryDirectory(prefix="fcfd_dropin_") as td:
        for d in ("in", "res", "out"):
            os.makedirs(os.path.join(td, d))
        with open(os.path.join(td, "case.txt"), "w") as fh:
            fh.write(case.to_reference_input(1))
        np.ascontiguousarray(ref["phi0"], dtype="<f8").reshape(-1).tofile(os.path.join(td, "in", "phi"))
        print(f"n2_dropin precision={precision} steps={nsteps}", flush=True)
        r = subprocess.run([DROPIN, "case.txt", "in", "res", str(nsteps), str(precision), repr(case.ppe_tol)],
                           cwd=td, env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, f"n2_dropin failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}"
        got = {f: np.fromfile(os.path.join(td, "res", f), dtype="<f8").reshape(case.shape)
               for f in ("phi", "u", "v", "p", "nu", "nv")}
        rows = np.loadtxt(os.path.join(td, "res", "mass.txt"), ndmin=2)
    assert rows.shape == (nsteps, 3)
    assert np.array_equal(rows[:, 0], np.arange(1, nsteps + 1))
    g = case.ghc
    worst = {}
    for f, a in got.items():
        b = ref[f"step{nsteps}/{f}"]
        worst[f] = float(np.abs(a[g:-g, g:-g] - b[g:-g, g:-g]).max() / np.abs(b[g:-g, g:-g]).max())
    dm = float(np.max(np.abs(rows[:, 1] - ref["mass"]) / ref["mass"]))
    print({k: f"{v:.2e}" for k, v in worst.items()}, f"mass {dm:.1e}")
    assert max(worst.values()) <= 1e-10, worst
    assert dm <= 1e-12
