"""Pin the FP64 reference gradient's forward (tests/adjoint_ref.py, torch on
the CPU) against the CPU oracle's level_set_rk3_solver, which is itself
bit-identical to the compiled reference: one SSP-RK3 step of the WENO5 flux
split on the same phi and cell velocity must agree to round-off.  Its
gradient is then torch.autograd of that forward."""
import numpy as np
import pytest

from cases import bubble, case3d, node_velocities
from oracle.oracle import Oracle

torch = pytest.importorskip("torch")


def test_reference_step_matches_oracle():
    import adjoint_ref as R
    case = case3d(14)
    f = {"phi": bubble(case, 0.3, 1), **node_velocities(case, 3)}
    o = Oracle(case)
    for k, v in f.items():
        o.set(k, v)
    o.call("switch")
    o.call("level_set_rk3_solver")      # the reference advects with nvel%old
    g = case.ghc
    inter = lambda a: np.ascontiguousarray(np.asarray(a).reshape(case.shape)[g:-g, g:-g, g:-g])
    vel = tuple(torch.from_numpy(inter(f[k])) for k in ("nu", "nv", "nw"))
    out = R.rk3_step(torch.from_numpy(inter(f["phi"])), vel, (case.dx, case.dy, case.dz), case.dt).numpy()
    ref = inter(o.get("phi"))
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print("max rel diff", rel)
    assert rel <= 1e-13


def test_reference_gradient_is_linear_in_the_seed():
    """d/dphi0 <w, phi_T> is linear in w (a property autograd must show)."""
    import adjoint_ref as R
    rng = np.random.default_rng(1)
    n = (6, 5, 7)
    phi = torch.from_numpy(rng.standard_normal(n[::-1]))
    u = tuple(torch.from_numpy(rng.standard_normal(n[::-1]) * 0.3) for _ in range(3))
    w1, w2 = (torch.from_numpy(rng.standard_normal(n[::-1])) for _ in range(2))

    def grad(w):
        p = phi.clone().requires_grad_(True)
        (R.advect(p, torch.tensor(1.0, dtype=torch.float64), u, (0.2, 0.2, 0.2), 0.02, 2, 1.0) * w).sum().backward()
        return p.grad

    g = grad(w1 + 2.0 * w2) - grad(w1) - 2.0 * grad(w2)
    assert float(g.abs().max()) <= 1e-12 * float(grad(w1).abs().max())


def test_device_deformation_field_is_the_ghost_copied_formula():
    """deformation_field_device (torch, clamped-index evaluation) equals the
    numpy field with the zero-gradient ghost copy, on the CPU."""
    from fortrancfd_amd.adjoint import deformation_field, deformation_field_device
    from fortrancfd_amd.types import Case
    case = Case(dim=3, nx=10, ny=7, nz=9, xend=1.0, yend=0.7, zend=0.9, cfl=0.3, name="adv")
    for a, b in zip(deformation_field(case), deformation_field_device(case, "cpu")):
        b = b.numpy().reshape(case.shape)
        assert np.abs(a - b).max() <= 4e-16 * np.abs(a).max()
        assert np.array_equal(b[0], b[case.ghc]) and np.array_equal(b[:, :, -1], b[:, :, -case.ghc - 1])


def test_ghost_fold_is_the_adjoint_of_the_ghost_copy():
    """<W, bc(x)> = <fold(W), x> for every ghosted x: fold adds each ghost's
    weight to the interior cell the zero-gradient copy takes it from."""
    import torch
    from fortrancfd_amd.adjoint import _fold_ghosts
    from fortrancfd_amd.types import Case, bc_copy
    for case in (Case(dim=3, nx=5, ny=4, nz=6, name="t"), Case(dim=2, nx=5, ny=7, nz=1, name="t")):
        rng = np.random.default_rng(1)
        W = rng.standard_normal(case.shape)
        fW = _fold_ghosts(torch.from_numpy(W), case.ghc).numpy()
        assert not fW[~_interior_mask(case)].any()
        for _ in range(3):
            x = rng.standard_normal(case.shape)
            lhs = float((W * bc_copy(case, x.copy())).sum())
            rhs = float((fW * x).sum())
            assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)


def _interior_mask(case):
    m = np.zeros(case.shape, dtype=bool)
    g = case.ghc
    if case.dim == 3:
        m[g:-g, g:-g, g:-g] = True
    else:
        m[g:-g, g:-g] = True
    return m
