"""Operation table shared by the oracle-vs-reference tests and the golden
fixture generator: for each reference operation, the inputs it needs, the
oracle calls that restate it and the outputs compared."""
from __future__ import annotations

from cases import bubble, node_velocities, seeded, vel_names, velocities


def _olds(case, phi):
    from oracle.oracle import Oracle
    o = Oracle(case)
    o.set("phi", phi)
    o.call("rho_mu")
    o.call("curv")
    out = {f + "_old": o.get(f) for f in ["rho", "mu", "delta", "curv", "nx", "ny"] + (["nz"] if case.dim == 3 else [])}
    out["phi_old"] = phi
    return out


def ops(case):
    """yield (op, inputs, scalars, oracle_calls, outputs)"""
    vn = vel_names(case)
    phi = bubble(case, 0.3, 1)
    V = velocities(case, 2)
    NV = node_velocities(case, 3)
    nf = ["nx", "ny", "nxx", "nyy", "nxy"] + (["nz", "nzz", "nxz", "nyz"] if case.dim == 3 else [])
    rho = seeded(case, 5) ** 2 + 0.001
    cf = ["coe_r", "coe_l", "coe_f", "coe_b", "coe_c"] + (["coe_u", "coe_d"] if case.dim == 3 else [])
    yield "bc_phi", {"phi": phi}, {}, [("bc", 0)], ["phi"]
    yield "velbc", dict(V), {}, [("velbc",)], vn
    yield "rho_mu", {"phi": phi}, {}, [("rho_mu",)], ["heavy", "delta", "sign", "rho", "mu"]
    yield "ls_mv", {"phi": phi}, {}, [("ls_mv",)], ["rho", "_mass", "_vol"]
    yield "ls_adv", {"phi": phi, **NV}, {}, [("switch",), ("level_set_rk3_solver",)], ["phi"]
    yield "redis1", {"phi": phi}, {}, [("level_set_rk3_redis", 1)], ["phi"]
    yield "redis0", {"phi": phi}, {}, [("level_set_rk3_redis", 0)], ["phi"]
    yield "mpls", {"phi": phi}, {"imass": 0.5}, [("mass_preserving_level_set",)], ["phi", "_mass"]
    yield "surface_norms", {"phi": phi}, {}, [("surface_norms",)], nf
    nd = ["nx", "ny", "nxx", "nyy"] + (["nz", "nzz"] if case.dim == 3 else [])
    yield "uccd", {"phi": phi, **NV}, {}, [("switch",), ("uccd",)], nd
    yield "curv", {"phi": phi}, {}, [("curv",)], nf + ["curv"]
    yield "ppe_init", {"rho": rho}, {}, [("ppe_sor_init",)], cf
    yield "projection", {"rho": rho, "p": seeded(case, 9), **V}, {}, [("ns_momentum_correction",)], vn
    yield "node_vel", dict(V), {}, [("node_vel",)], ["nu", "nv"] + (["nw"] if case.dim == 3 else [])
    inp = {**V, **{k + "_old": velocities(case, 4)[k] for k in vn}, **_olds(case, phi)}
    for k in vn:
        inp["src_" + k + "_old"] = seeded(case, 11 + len(k), 0.1)
    yield ("ns_setup", inp, {}, [("ns_linearize",), ("ns_ab_setup",), ("ns_ab_predictor",)],
           vn + ["src_" + k for k in vn])
    yield ("ppe_solve", {"rho": rho, **V}, {}, [("ppe_sor_init",), ("ppe_sor_solver", 0.1 ** case.p_tol_exp)],
           ["p"] + vn + ["_piter"])
    yield ("init", {"phi": phi}, {}, [("problem_init", 1)],
           ["phi", "rho", "curv", "_mass"] + ["src_" + k for k in vn])
