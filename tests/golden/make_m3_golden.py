"""Write tests/golden/m3_ref.npz: outputs of the compiled MPLS_RisingBubble
reference (oracle/_ref/refm3, built by oracle/build_ref.sh from the
reference's own sources) on the seeded two-bubble state of tests/m3cases.py
(12 x 16 x 16 cells): one call each of AMURHO, LS_CCD, LS_sorce, LS_RK3_SOLVER,
LS_REDISTANCE, MPLS and SOLVE_LS, then the post-initial-condition part of
SETUP and three main-loop steps with the reference's own SOR pressure
solve.  Data only (inputs are regenerated from the seeds).
Run from the repository root:  python tests/golden/make_m3_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from m3cases import GOLDEN_OPS, golden_inputs, m3_case  # noqa: E402
from oracle.m3 import FIELDS, harness_available, run_reference  # noqa: E402

NSTEPS = 3


def main():
    if not harness_available():
        raise SystemExit("reference harness not built: run oracle/build_ref.sh")
    case = m3_case(12)
    blob = {}
    inp, state = golden_inputs(case)
    for op, outs in GOLDEN_OPS.items():
        ref = run_reference(case, op, inp, **state)
        for n in outs:
            blob[f"{op}/{n}"] = ref[n]
    ref = run_reference(case, "setup_tail", golden_inputs(case, bubbles_only=True)[0], rec_mass=0, iter=0)
    for n in ("phi", "us", "curv"):
        blob[f"setup_tail/{n}"] = ref[n]
    for n in range(1, NSTEPS + 1):
        sc = ref["_scalars"]
        ref = run_reference(case, "step", {k: ref[k] for k in FIELDS}, b1_im=sc["b1_im"], b2_im=sc["b2_im"],
                            imass_ls=sc["imass_ls"], ivol_ls=sc["ivol_ls"], time=sc["time"], iter=sc["iter"],
                            rec_mass=sc["rec_mass"])
        blob[f"step{n}/mass_ls"] = np.float64(ref["_scalars"]["mass_ls"])
    for f in ("phi", "phi1", "u", "v", "w", "p"):
        blob[f"step{NSTEPS}/{f}"] = ref[f]
    out = os.path.join(HERE, "m3_ref.npz")
    np.savez_compressed(out, **blob)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
