"""Write the golden fixtures tests/golden/ref_<dim>d.npz from the compiled
Fortran reference harness (oracle/_ref, built by oracle/build_ref.sh).

Each fixture holds, per reference operation of tests/refops.py, the seeded
inputs and the reference's outputs (data only: float64 arrays and scalars).
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import case2d, case3d  # noqa: E402
from oracle import refrun  # noqa: E402
from refops import ops  # noqa: E402

GOLDEN_CASES = {2: lambda: case2d(16, p_tol_exp=8.0), 3: lambda: case3d(8, p_tol_exp=8.0)}


def main():
    for dim, mk in GOLDEN_CASES.items():
        case = mk()
        if not refrun.available(dim):
            raise SystemExit("reference harness not built: run oracle/build_ref.sh")
        blob = {}
        for op, inputs, scalars, calls, outs in ops(case):
            ref = refrun.run(case, op, inputs, imass=scalars.get("imass", 0.0))
            for k, v in inputs.items():
                blob[f"{op}/in/{k}"] = v
            for f in outs:
                blob[f"{op}/out/{f}"] = np.asarray(ref[f])
        path = os.path.join(HERE, f"ref_{dim}d.npz")
        np.savez_compressed(path, **blob)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
