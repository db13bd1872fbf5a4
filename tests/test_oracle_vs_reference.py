"""Pin the CPU oracle against the compiled Fortran reference (oracle/_ref).

The reference harness runs the reference's own subroutines on seeded inputs
(single-threaded, so its reductions and its in-place relaxation are
sequential); the oracle must agree bit for bit.  Skipped where the harness is
not built (it needs the reference checkout and flang)."""
import numpy as np
import pytest

from cases import bubble, case2d, case3d, max_rel, vel_names
from oracle import refrun
from oracle.oracle import Oracle
from refops import ops

CASES = [("2d", lambda: case2d(16, p_tol_exp=8.0)), ("3d", lambda: case3d(8, p_tol_exp=8.0))]


def _run_oracle(case, inputs, scalars, calls):
    Oracle.threads(1)
    o = Oracle(case)
    for k, v in inputs.items():
        o.set(k, v)
    o.set_imass(scalars.get("imass", 0.0))
    o.set_time(0.0, scalars.get("it", 0))
    for c in calls:
        o.call(*c)
    return o


def _value(o, name):
    d = o.diag()
    return {"_mass": d["mass"], "_vol": d["vol"], "_piter": d["ppe_cycles"]}[name] if name.startswith("_") else o.get(name)


@pytest.mark.parametrize("tag,mk", CASES)
def test_every_subroutine_bitwise(tag, mk):
    case = mk()
    if not refrun.available(case.dim):
        pytest.skip("reference harness not built (oracle/build_ref.sh)")
    for op, inputs, scalars, calls, outs in ops(case):
        ref = refrun.run(case, op, inputs, imass=scalars.get("imass", 0.0))
        o = _run_oracle(case, inputs, scalars, calls)
        for f in outs:
            a, b = _value(o, f), ref[f]
            if f.startswith("_"):
                assert a == b, (op, f, a, b)
            else:
                assert np.array_equal(a, b), (op, f, max_rel(a, b))


@pytest.mark.parametrize("tag,mk", CASES)
def test_time_steps_bitwise(tag, mk):
    """init + steps 4 and 5 (step 5 reinitialises) of the main loop, with the
    reference's relaxation pressure solver run sequentially."""
    case = mk(); case.p_tol_exp = 10.0
    if not refrun.available(case.dim):
        pytest.skip("reference harness not built (oracle/build_ref.sh)")
    Oracle.threads(1)
    phi = bubble(case, 0.0, 1)
    st = refrun.run(case, "init_nore", {"phi": phi})
    names = refrun.HARNESS_FIELDS_3D if case.dim == 3 else refrun.HARNESS_FIELDS_2D
    o = Oracle(case)
    o.set("phi", phi)
    o.call("problem_init", 0)
    assert o.diag()["imass"] == st["_imass"]
    o.set_time(0.0, 3)
    t = 0.0
    for it in (4, 5):
        st = refrun.run(case, "step", {k: st[k] for k in names}, imass=st["_imass"], time=t, it=it - 1)
        t = st["_time"]
        o.call("step", 0)
        for f in ["phi", "p", "rho", "curv", "nu", "src_u"] + vel_names(case):
            assert np.array_equal(o.get(f), st[f]), (it, f, max_rel(o.get(f), st[f]))
        assert o.diag()["ppe_cycles"] == st["_piter"]
        assert o.diag()["mass"] == st["_mass"]
