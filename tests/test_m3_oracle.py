"""The MPLS_RisingBubble oracle (oracle/m3_oracle.c) pinned bit for bit
against the compiled reference: against the committed fixture
tests/golden/m3_ref.npz (written by tests/golden/make_m3_golden.py from
oracle/_ref/refm3), and, where the harness is built, against live runs of
every restated subroutine and of the setup plus five main-loop steps.
Single-threaded on both sides (the reference's SOR updates p in place
inside an OpenMP loop)."""
import os

import numpy as np
import pytest

from m3cases import GOLDEN_OPS, bubbles, golden_inputs, m3_case
from oracle.m3 import FIELDS, M3Oracle, harness_available, run_reference
from oracle.oracle import Oracle

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "m3_ref.npz")

ORACLE_CALL = {
    "amurho": lambda o: o.call("amurho"),
    "ls_ccd": lambda o: o.call("ls_ccd"),
    "ls_source": lambda o: o.call("ls_source"),
    "ls_rk3": lambda o: o.call("ls_rk3_solver"),
    "redistance": lambda o: o.call("ls_redistance"),
    "mpls1": lambda o: o.call("mpls", o.params()["b1_im"]),
    "solve_ls": lambda o: o.call("solve_ls"),
    "second_order_upwind": lambda o: o.call("second_order_upwind"),
}


def _oracle(case, inp, state):
    Oracle.threads(1)
    o = M3Oracle(case, **state)
    for k, v in inp.items():
        o.set(k, v)
    return o


def _steps(case, n):
    o = _oracle(case, golden_inputs(case, bubbles_only=True)[0], dict(rec_mass=0, iter=0))
    o.call("setup_tail", 0)
    out = [o.get("phi").copy()]
    masses = []
    for _ in range(n):
        o.call("step", 0)
        masses.append(o.diag()["mass_ls"])
    return o, masses


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="tests/golden/m3_ref.npz not generated")
@pytest.mark.parametrize("op", sorted(GOLDEN_OPS))
def test_m3_oracle_matches_the_fixture(op):
    case = m3_case(12)
    ref = np.load(FIXTURE)
    inp, state = golden_inputs(case)
    o = _oracle(case, inp, state)
    ORACLE_CALL[op](o)
    for n in GOLDEN_OPS[op]:
        assert np.array_equal(o.get(n), ref[f"{op}/{n}"]), (op, n)


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="tests/golden/m3_ref.npz not generated")
def test_m3_oracle_setup_and_steps_match_the_fixture():
    case = m3_case(12)
    ref = np.load(FIXTURE)
    o = _oracle(case, golden_inputs(case, bubbles_only=True)[0], dict(rec_mass=0, iter=0))
    o.call("setup_tail", 0)
    for n in ("phi", "us", "curv"):
        assert np.array_equal(o.get(n), ref[f"setup_tail/{n}"]), n
    for step in range(1, 4):
        o.call("step", 0)
        assert o.diag()["mass_ls"] == float(ref[f"step{step}/mass_ls"]), step
    for n in ("phi", "phi1", "u", "v", "w", "p"):
        assert np.array_equal(o.get(n), ref[f"step3/{n}"]), n


@pytest.mark.skipif(not harness_available(), reason="oracle/_ref/refm3 not built")
@pytest.mark.parametrize("op,call,outs", [
    ("heavy_f", lambda o: o.call("heavy_f"), ("heavy", "delta")),
    ("gradfi", lambda o: o.call("gradfi"), ("grad", "sgn")),
    ("rhsp", lambda o: (o.call("adams_bashforth"), o.call("rhsp")), ("rhs", "u1", "w1")),
    ("solve_ns", lambda o: o.call("solve_ns", 0), ("u", "v", "w", "p", "us")),
    ("find_mass", lambda o: o.call("find_mass"), ("rho", "heavy")),
])
def test_m3_oracle_matches_the_compiled_reference(op, call, outs):
    case = m3_case(12)
    inp, state = golden_inputs(case)
    if op == "rhsp":
        for k in ("us", "vs", "ws"):
            inp[k] = 0.1 * inp[k[0]]
            inp[k + "_old"] = 0.2 * inp[k[0]]
        ref = run_reference(case, "none", inp, **state)
        o = _oracle(case, inp, state)
        o.call("adams_bashforth")
        for k in ("u1", "v1", "w1"):
            inp[k] = o.get(k)
        ref = run_reference(case, "rhsp", inp, **state)
        o.call("rhsp")
    else:
        if op == "find_mass":
            state = dict(state, rec_mass=0)
        ref = run_reference(case, op, inp, **state)
        o = _oracle(case, inp, state)
        call(o)
    for n in outs:
        assert np.array_equal(o.get(n), ref[n]), (op, n)
    if op == "find_mass":
        p = o.params()
        assert p["b1_im"] == ref["_scalars"]["b1_im"] and p["b2_im"] == ref["_scalars"]["b2_im"]
        assert p["imass_ls"] == ref["_scalars"]["imass_ls"]


@pytest.mark.skipif(not harness_available(), reason="oracle/_ref/refm3 not built")
def test_m3_oracle_five_steps_match_the_compiled_reference():
    """setup + 5 main-loop steps (the fifth maintains both bubbles' level
    sets), the reference's SOR pressure solve on both sides: every field
    bit for bit after every step."""
    case = m3_case(12)
    inp = bubbles(case)
    ref = run_reference(case, "setup_tail", inp, rec_mass=0, iter=0)
    o = _oracle(case, inp, dict(rec_mass=0, iter=0))
    o.call("setup_tail", 0)
    for step in range(1, 6):
        sc = ref["_scalars"]
        ref = run_reference(case, "step", {k: ref[k] for k in FIELDS}, b1_im=sc["b1_im"], b2_im=sc["b2_im"],
                            imass_ls=sc["imass_ls"], ivol_ls=sc["ivol_ls"], time=sc["time"], iter=sc["iter"],
                            rec_mass=sc["rec_mass"])
        o.call("step", 0)
        bad = [f for f in FIELDS if not np.array_equal(o.get(f), ref[f])]
        assert not bad, (step, bad)
        assert o.diag()["mass_ls"] == ref["_scalars"]["mass_ls"]


def test_m3_converged_pressure_solve_is_the_sor_fixed_point():
    """the oracle's direct solve of PRESS_POISSON's operator (the fixed point
    of the reference's SOR, which the GPU path is compared with) against the
    SOR run far past the reference's tolerance."""
    case = m3_case(12, tol_p=1e-13)
    inp, state = golden_inputs(case)
    # no flow through the walls (as bcu / bcv / bcw leave it): the source
    # div(u_old / dt - U1) then sums to zero and the system is compatible
    inp["u_old"][:, :, 2] = 0.0; inp["u_old"][:, :, case.nx + 2] = 0.0
    inp["v_old"][:, 2, :] = 0.0; inp["v_old"][:, case.ny + 2, :] = 0.0
    inp["w_old"][2] = 0.0; inp["w_old"][case.nz + 2] = 0.0
    o = _oracle(case, inp, state)
    o.call("second_order_upwind")   # US, VS, WS as a step forms them (zero on the wall faces)
    o.call("adams_bashforth")
    o.call("rhsp")
    o.call("press_poisson")
    p_sor = case.interior(o.get("p")).copy()
    o.set("p", np.zeros(case.shape))
    o.call("press_exact")
    p_dir = case.interior(o.get("p"))
    r = float(np.abs(p_dir - p_sor).max() / np.abs(p_sor).max())
    print("direct vs SOR to 1e-13", r)
    assert r <= 1e-9
