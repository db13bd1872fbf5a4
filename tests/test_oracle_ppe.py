"""The oracle's converged pressure solve against a library direct solve.

The reference's relaxation (ns_ppe_solver.f90:67-172) updates p in place
inside an OpenMP loop, so its iterate is not reproducible; what parity can
rest on is the fixed point it converges to: A p = src - mean(src), mean(p) = 0
with A the ppe_sor_init operator.  The oracle solves that system with a band
Cholesky factor of the pinned operator (and, for systems too large to factor,
with a projected Jacobi-PCG).  Both are pinned here against SciPy's sparse LU
of the same operator, assembled from the oracle's coefficient arrays (which
are themselves bitwise equal to the compiled reference's ppe_sor_init,
test_oracle_vs_reference.py), and by the residual of their solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from cases import bubble, case2d, case3d, max_rel, velocities
from fortrancfd_amd.types import Case
from oracle.oracle import Oracle, lib


def _odd3d(**kw):
    return Case(dim=3, nx=20, ny=18, nz=37, xend=1.0, yend=0.9, zend=1.85, gvec=(0.0, 0.0, -1.0),
                name="t3d", **kw)


CASES = [("2d", lambda: case2d(32)), ("3d", lambda: case3d(16)), ("3d_odd", _odd3d)]


def _system(case):
    """oracle state with rho of a bubble (ratio 1000) and a seeded velocity,
    after ppe_sor_init and the PPE source."""
    Oracle.threads(1)
    tmp = Oracle(case)
    tmp.set("phi", bubble(case, 0.0, 1))
    tmp.call("rho_mu")
    o = Oracle(case)
    o.set("rho", tmp.get("rho"))
    for k, v in velocities(case, 6).items():
        o.set(k, v)
    o.call("velbc")
    o.call("ppe_sor_init")
    return o


def _scipy_solution(case, o, src):
    ci = case.interior
    C = {k: ci(o.get("coe_" + k)) for k in ("r", "l", "f", "b", "u", "d", "c")}
    shp = C["c"].shape
    n = int(np.prod(shp))
    idx = np.arange(n).reshape(shp)
    offs = {"r": (0, 0, 1), "l": (0, 0, -1), "f": (0, 1, 0), "b": (0, -1, 0), "u": (1, 0, 0), "d": (-1, 0, 0)}
    if case.dim == 2:
        offs = {k: v[1:] for k, v in offs.items() if k in "rlfb"}
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [C["c"].ravel()]
    for k, off in offs.items():
        nz = np.nonzero(C[k] != 0)
        tgt = tuple(a + d for a, d in zip(nz, off))
        rows.append(idx[nz]); cols.append(idx[tgt]); vals.append(C[k][nz])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    f = ci(src).ravel() - ci(src).mean()
    x = np.zeros(n)
    x[1:] = spl.spsolve(A[1:, 1:].tocsc(), f[1:])
    x -= x.mean()
    return x.reshape(shp), A, f


@pytest.mark.parametrize("tag,mk", CASES)
@pytest.mark.parametrize("band", [True, False], ids=["direct", "pcg"])
def test_converged_solve_matches_sparse_lu(tag, mk, band):
    case = mk()
    o = _system(case)
    lib().orc_set_band_max(400 * 1000 * 1000 if band else 0)
    try:
        o.call("ppe_exact_solver", 1e-13)
    finally:
        lib().orc_set_band_max(400 * 1000 * 1000)
    src = o.get("ppe_src")
    ref, A, f = _scipy_solution(case, o, src)
    # the oracle projected the velocity afterwards; p is the solve's output
    p = case.interior(o.get("p"))
    r = max_rel(p, ref)
    print(tag, "band" if band else "pcg", "p vs sparse LU", r)
    assert r <= (1e-11 if band else 1e-10), r
    assert abs(p.mean()) <= 1e-15 * np.abs(p).max()
    res = np.linalg.norm(f - A @ p.ravel()) / np.linalg.norm(f)
    ref_res = np.linalg.norm(f - A @ ref.ravel()) / np.linalg.norm(f)
    print("rel residual: oracle", res, "sparse LU", ref_res)
    assert res <= max(10 * ref_res, 1e-13 if band else 1e-12), (res, ref_res)


def test_solve_rhs_alone_leaves_the_velocity():
    """orc_ppe_solve_rhs solves the given source with no projection."""
    case = case3d(16)
    o = _system(case)
    rng = np.random.default_rng(5)
    src = rng.uniform(-1, 1, case.shape)
    o.set("ppe_src", src)
    u0 = o.get("u").copy()
    o.call("ppe_solve_rhs", 1e-13)
    assert np.array_equal(o.get("u"), u0)
    ref, _, _ = _scipy_solution(case, o, src)
    assert max_rel(case.interior(o.get("p")), ref) <= 1e-11
