"""Adjoint of the WENO5 SSP-RK3 level-set advection (fcfd_ls_adv_adjoint_step
and the torch.autograd.Function of fortrancfd_amd.adjoint).

* one step: the adjoint of the level set and of the velocity against
  torch.autograd of the independent restatement (tests/adjoint_ref.py) on
  the same state, and <w, dS(phi) v> by central differences of the device
  forward step against <dS(phi)^T w, v> (the dot-product test of a
  discrete adjoint).  The perturbations v are smooth fields: WENO-JS with
  EPS = 1e-13 responds to white noise of size 1e-6 far outside its linear
  range, so a rough v measures the curvature of the weights, not the
  derivative;
* a run with checkpointing: dJ/dphi0 along a random direction and dJ/dA
  against central finite differences with eps = 1e-6, the acceptance bound
  1e-6 relative;
* the gradient fields against an independent FP64 reference gradient:
  torch.autograd of the same forward restated on the CPU
  (tests/adjoint_ref.py), whose forward is checked against the device first;
* the velocity scale (a step with s (NU, NV, NW) equals a step with the
  scaled fields, bit for bit when the scaling is exact) and the FP32
  forward (fcfd_level_set_rk3_solver_f32) against the FP64 path and against
  the restatement of tests/adjoint_ref.py.
"""
import math

import numpy as np
import pytest

from cases import _smooth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(n=(20, 16, 18), **kw):
    from fortrancfd_amd.types import Case
    nx, ny, nz = n
    return Case(dim=3, nx=nx, ny=ny, nz=nz, xend=1.0, yend=ny / nx, zend=nz / nx, cfl=0.3, name="adv", **kw)


def _phi0(case, r=0.15, c=(0.35, 0.35, 0.35)):
    from fortrancfd_amd.types import bc_copy, bubble_phi
    return bc_copy(case, bubble_phi(case, r, c))


def _solver(case):
    from fortrancfd_amd import Solver
    return Solver(case)


def _interior_random(case, rng):
    a = np.zeros(case.shape)
    g = case.ghc
    a[g:-g, g:-g, g:-g] = rng.standard_normal((case.nz, case.ny, case.nx))
    return a


def _step(s, case, phi, vel):
    for nm, v in zip(("nu", "nv", "nw"), vel):
        s.push(nm, v)
    s.push("phi", phi)
    s.level_set_rk3_solver()
    return s.pull("phi")


def test_adjoint_step_dot_product():
    from fortrancfd_amd.adjoint import deformation_field
    case = _case()
    s = _solver(case)
    rng = np.random.default_rng(7)
    phi = _phi0(case)
    vel = [0.8 * u for u in deformation_field(case)]
    v = _smooth(case, rng, 1.0)
    q = [_smooth(case, rng, 1.0) for _ in range(3)]
    w = _interior_random(case, rng)
    g = case.ghc
    inner = lambda a, b: float((a[g:-g, g:-g, g:-g] * b[g:-g, g:-g, g:-g]).sum())
    eps = 1e-6
    # directional derivatives of one forward step by central differences
    dphi = (_step(s, case, phi + eps * v, vel) - _step(s, case, phi - eps * v, vel)) / (2 * eps)
    dvel = (_step(s, case, phi, [a + eps * b for a, b in zip(vel, q)]) -
            _step(s, case, phi, [a - eps * b for a, b in zip(vel, q)])) / (2 * eps)
    # adjoint at the same state
    for nm, a in zip(("nu", "nv", "nw"), vel):
        s.push(nm, a)
    dev = torch.device("cuda")
    ph_t = torch.from_numpy(phi.reshape(-1).copy()).to(dev)
    lam = torch.from_numpy(w.reshape(-1).copy()).to(dev)
    bars = [torch.zeros(case.size, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    s.ls_adv_adjoint_step(ph_t.data_ptr(), lam.data_ptr(), *(b.data_ptr() for b in bars))
    s.sync()
    lam_n = lam.cpu().numpy().reshape(case.shape)
    bars_n = [b.cpu().numpy().reshape(case.shape) for b in bars]
    lhs, rhs = inner(w, dphi), inner(lam_n, v)
    print("phi: <w, dS v> =", lhs, " <dS^T w, v> =", rhs)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs)
    lhs_u = inner(w, dvel)
    rhs_u = sum(inner(b, c) for b, c in zip(bars_n, q))
    print("vel: <w, dS q> =", lhs_u, " <dS_u^T w, q> =", rhs_u)
    assert abs(lhs_u - rhs_u) <= 1e-7 * abs(lhs_u)


def test_adjoint_step_against_autograd():
    import adjoint_ref as R
    from fortrancfd_amd.adjoint import deformation_field
    case = _case()
    g = case.ghc
    inter = lambda a: np.ascontiguousarray(np.asarray(a).reshape(case.shape)[g:-g, g:-g, g:-g])
    phi = _phi0(case)
    vel = [0.8 * u for u in deformation_field(case)]
    w = _interior_random(case, np.random.default_rng(7))
    s = _solver(case)
    for nm, a in zip(("nu", "nv", "nw"), vel):
        s.push(nm, a)
    dev = torch.device("cuda")
    ph_t = torch.from_numpy(phi.reshape(-1).copy()).to(dev)
    lam = torch.from_numpy(w.reshape(-1).copy()).to(dev)
    bars = [torch.zeros(case.size, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    s.ls_adv_adjoint_step(ph_t.data_ptr(), lam.data_ptr(), *(b.data_ptr() for b in bars))
    s.sync()
    p = torch.from_numpy(inter(phi)).requires_grad_(True)
    u = [torch.from_numpy(inter(a)).requires_grad_(True) for a in vel]
    out = R.rk3_step(p, tuple(u), (case.dx, case.dy, case.dz), case.dt)
    (out * torch.from_numpy(inter(w))).sum().backward()
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    r = [rel(inter(lam.cpu().numpy()), p.grad.numpy())]
    r += [rel(inter(b.cpu().numpy()), uu.grad.numpy()) for b, uu in zip(bars, u)]
    print("adjoint vs autograd (phi, u, v, w):", r)
    assert max(r) <= 1e-12
    # ghost entries of lambda are left as given
    m = np.ones(case.shape, dtype=bool)
    m[g:-g, g:-g, g:-g] = False
    assert np.array_equal(lam.cpu().numpy().reshape(case.shape)[m], w[m])


def _loss_and_grad(case, phi0, A, target, nsteps, period, every):
    from fortrancfd_amd.adjoint import advect, l2_loss
    s = _solver(case)
    dev = torch.device("cuda")
    p0 = torch.from_numpy(phi0.reshape(-1).copy()).to(dev).requires_grad_(True)
    a = torch.tensor(A, dtype=torch.float64, device=dev, requires_grad=True)
    J = l2_loss(advect(p0, a, s, nsteps, period, every), target, case)
    J.backward()
    return float(J), p0.grad.cpu().numpy().reshape(case.shape), float(a.grad)


def _loss(case, phi0, A, target, nsteps, period):
    from fortrancfd_amd.adjoint import advect, l2_loss
    s = _solver(case)
    dev = torch.device("cuda")
    with torch.no_grad():
        p0 = torch.from_numpy(phi0.reshape(-1).copy()).to(dev)
        a = torch.tensor(A, dtype=torch.float64, device=dev)
        return float(l2_loss(advect(p0, a, s, nsteps, period, nsteps), target, case))


def test_gradient_against_finite_differences():
    from fortrancfd_amd.adjoint import advect
    case = _case((24, 24, 24))
    nsteps, period, every = 30, 30 * case.dt, 7     # uneven last segment (30 = 4*7 + 2)
    phi0 = _phi0(case)
    dev = torch.device("cuda")
    with torch.no_grad():
        target = advect(torch.from_numpy(phi0.reshape(-1).copy()).to(dev),
                        torch.tensor(1.05, dtype=torch.float64, device=dev), _solver(case), nsteps, period, nsteps)
    J, gphi, gA = _loss_and_grad(case, phi0, 1.0, target, nsteps, period, every)
    rng = np.random.default_rng(3)
    v = _smooth(case, rng, 1.0)
    eps = 1e-6
    fd_phi = (_loss(case, phi0 + eps * v, 1.0, target, nsteps, period) -
              _loss(case, phi0 - eps * v, 1.0, target, nsteps, period)) / (2 * eps)
    ad_phi = float((gphi * v).sum())   # gphi is zero on ghosts: v's ghosts do not count
    fd_A = (_loss(case, phi0, 1.0 + eps, target, nsteps, period) -
            _loss(case, phi0, 1.0 - eps, target, nsteps, period)) / (2 * eps)
    print("J", J, "phi:", ad_phi, fd_phi, "A:", gA, fd_A)
    assert abs(ad_phi - fd_phi) <= 1e-6 * abs(fd_phi)
    assert abs(gA - fd_A) <= 1e-6 * abs(fd_A)
    # ghost cells of phi0 do not enter the advection
    g = case.ghc
    m = np.ones(case.shape, dtype=bool)
    m[g:-g, g:-g, g:-g] = False
    assert not gphi[m].any()


def test_gradient_against_autograd_reference():
    import adjoint_ref as R
    from fortrancfd_amd.adjoint import advect, deformation_field
    case = _case((12, 10, 14))
    nsteps, period, every = 6, 6 * case.dt, 4
    phi0 = _phi0(case, r=0.2, c=(0.45, 0.4, 0.5))
    g = case.ghc
    inter = lambda a: np.ascontiguousarray(a.reshape(case.shape)[g:-g, g:-g, g:-g])
    uhat = tuple(torch.from_numpy(inter(u)) for u in deformation_field(case))
    h = (case.dx, case.dy, case.dz)
    rng = np.random.default_rng(9)
    wgt = rng.standard_normal((case.nz, case.ny, case.nx))
    # reference forward + autograd (CPU, FP64)
    p0 = torch.from_numpy(inter(phi0)).requires_grad_(True)
    A = torch.tensor(0.9, dtype=torch.float64, requires_grad=True)
    pT = R.advect(p0, A, uhat, h, case.dt, nsteps, period)
    (pT * torch.from_numpy(wgt)).sum().backward()
    # device forward + hand-written adjoint with the same linear loss
    dev = torch.device("cuda")
    q0 = torch.from_numpy(phi0.reshape(-1).copy()).to(dev).requires_grad_(True)
    B = torch.tensor(0.9, dtype=torch.float64, device=dev, requires_grad=True)
    qT = advect(q0, B, _solver(case), nsteps, period, every)
    W = np.zeros(case.shape)
    W[g:-g, g:-g, g:-g] = wgt
    (qT * torch.from_numpy(W.reshape(-1)).to(dev)).sum().backward()
    fwd = np.abs(inter(qT.detach().cpu().numpy()) - pT.detach().numpy()).max() / np.abs(pT.detach().numpy()).max()
    gp = inter(q0.grad.cpu().numpy())
    rel = np.abs(gp - p0.grad.numpy()).max() / np.abs(p0.grad.numpy()).max()
    relA = abs(float(B.grad) - float(A.grad)) / abs(float(A.grad))
    print("forward", fwd, "grad phi", rel, "grad A", relA)
    assert fwd <= 1e-13
    assert rel <= 1e-10
    assert relA <= 1e-10


def test_gradient_with_weights_on_the_ghosts():
    """A loss that weights every entry of the ghosted phi_T, ghosts included:
    the ghosts of phi_T are zero-gradient copies of the boundary cells, so
    their weights reach dJ/dphi0 and dJ/dA through those cells.  Against
    torch.autograd of the restatement with the same copy (replicate padding)
    in its forward; dJ/dphi0 is zero on phi0's ghosts, which the advection
    never reads."""
    import adjoint_ref as R
    from fortrancfd_amd.adjoint import advect, deformation_field
    case = _case((12, 10, 14))
    nsteps, period, every = 5, 5 * case.dt, 3
    phi0 = _phi0(case, r=0.2, c=(0.3, 0.4, 0.5))
    g = case.ghc
    inter = lambda a: np.ascontiguousarray(a.reshape(case.shape)[g:-g, g:-g, g:-g])
    uhat = tuple(torch.from_numpy(inter(u)) for u in deformation_field(case))
    h = (case.dx, case.dy, case.dz)
    W = np.random.default_rng(11).standard_normal(case.shape)
    p0 = torch.from_numpy(inter(phi0)).requires_grad_(True)
    A = torch.tensor(1.1, dtype=torch.float64, requires_grad=True)
    pT = R.advect(p0, A, uhat, h, case.dt, nsteps, period)
    full = torch.nn.functional.pad(pT[None, None], (g,) * 6, mode="replicate")[0, 0]
    (full * torch.from_numpy(W)).sum().backward()
    dev = torch.device("cuda")
    q0 = torch.from_numpy(phi0.reshape(-1).copy()).to(dev).requires_grad_(True)
    B = torch.tensor(1.1, dtype=torch.float64, device=dev, requires_grad=True)
    qT = advect(q0, B, _solver(case), nsteps, period, every)
    fwd = np.abs(qT.detach().cpu().numpy().reshape(case.shape) - full.detach().numpy()).max()
    (qT * torch.from_numpy(W.reshape(-1)).to(dev)).sum().backward()
    gq = q0.grad.cpu().numpy().reshape(case.shape)
    rel = np.abs(inter(gq) - p0.grad.numpy()).max() / np.abs(p0.grad.numpy()).max()
    relA = abs(float(B.grad) - float(A.grad)) / abs(float(A.grad))
    m = np.ones(case.shape, dtype=bool)
    m[g:-g, g:-g, g:-g] = False
    print("forward", fwd, "grad phi", rel, "grad A", relA)
    assert fwd <= 1e-13 * np.abs(full.detach().numpy()).max()
    assert rel <= 1e-10 and relA <= 1e-10
    assert not gq[m].any()


def test_velocity_scale_is_a_factor_on_the_velocity():
    from fortrancfd_amd.adjoint import deformation_field
    case = _case()
    s = _solver(case)
    phi = _phi0(case)
    vel = deformation_field(case)
    ref = _step(s, case, phi, [0.7 * u for u in vel])
    s.set_velocity_scale(0.5)
    half = _step(s, case, phi, [1.4 * u for u in vel])    # 0.5 (1.4 u) and 0.7 u differ by rounding
    s.set_velocity_scale(-1.0)
    neg = _step(s, case, phi, [-0.7 * u for u in vel])     # exact: a sign flip
    s.set_velocity_scale(1.0)
    assert np.array_equal(neg, ref)
    assert np.abs(half - ref).max() <= 1e-14 * np.abs(ref).max()


def test_fp32_forward_against_fp64():
    import adjoint_ref as R
    from fortrancfd_amd.adjoint import advect, advect_f32, deformation_field
    case = _case()
    g = case.ghc
    inter = lambda a: np.asarray(a, dtype=np.float64).reshape(case.shape)[g:-g, g:-g, g:-g]
    s = _solver(case)
    dev = torch.device("cuda")
    phi0 = _phi0(case)
    nsteps, period = 12, 3.0
    p64 = advect(torch.from_numpy(phi0.reshape(-1).copy()).to(dev), torch.tensor(1.0, dtype=torch.float64, device=dev),
                 s, nsteps, period, nsteps).cpu().numpy()
    p32 = advect_f32(torch.from_numpy(phi0.reshape(-1).copy()).to(dev), 1.0, s, nsteps, period).cpu().numpy()
    assert p32.dtype == np.float32
    scale = np.abs(inter(p64)).max()
    err = np.abs(inter(p32) - inter(p64)).max() / scale
    # the restatement in FP32 on the CPU (same formulas, its own rounding)
    uhat = tuple(torch.from_numpy(np.ascontiguousarray(inter(u))).float() for u in deformation_field(case))
    pr = R.advect(torch.from_numpy(np.ascontiguousarray(inter(phi0))).float(), torch.tensor(1.0), uhat,
                  (case.dx, case.dy, case.dz), case.dt, nsteps, period).numpy()
    err_r = np.abs(inter(p32) - pr).max() / scale
    print("FP32 vs FP64 device:", err, " FP32 device vs FP32 restatement:", err_r)
    assert err <= 2e-6
    assert err_r <= 2e-6
    # ghosts hold the zero-gradient copy
    full = p32.reshape(case.shape)
    assert np.array_equal(full[:, :, 0], full[:, :, g])
    assert np.array_equal(full[:, :, -1], full[:, :, -g - 1])


def test_fp32_adjoint_step_against_fp64():
    """The FP32 adjoint of one FP32 step (fcfd_ls_adv_adjoint_step_f32) at a
    state representable in FP32, against the FP64 adjoint at the same state:
    the two differ by FP32 rounding of the same derivative."""
    from fortrancfd_amd.adjoint import deformation_field
    case = _case()
    s = _solver(case)
    rng = np.random.default_rng(7)
    phi = _phi0(case).astype(np.float32).astype(np.float64)
    vel = [(0.8 * u).astype(np.float32).astype(np.float64) for u in deformation_field(case)]
    w = _interior_random(case, rng).astype(np.float32).astype(np.float64)
    dev = torch.device("cuda")
    for nm, a in zip(("nu", "nv", "nw"), vel):
        s.push(nm, a)
    res = {}
    for dt_, nm in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        ph = torch.from_numpy(phi.reshape(-1).copy()).to(dev, dt_)
        lam = torch.from_numpy(w.reshape(-1).copy()).to(dev, dt_)
        bars = [torch.zeros(case.size, dtype=dt_, device=dev) for _ in range(3)]
        vt = [torch.from_numpy(a.reshape(-1).copy()).to(dev, dt_) for a in vel]
        torch.cuda.synchronize()
        if dt_ == torch.float64:
            s.ls_adv_adjoint_step(ph.data_ptr(), lam.data_ptr(), *(b.data_ptr() for b in bars))
        else:
            s.ls_adv_adjoint_step_f32(ph.data_ptr(), lam.data_ptr(), *(b.data_ptr() for b in bars),
                                      *(v.data_ptr() for v in vt))
        s.sync()
        res[nm] = [lam.double().cpu().numpy()] + [b.double().cpu().numpy() for b in bars]
    g = case.ghc
    inter = lambda a: a.reshape(case.shape)[g:-g, g:-g, g:-g]
    rel = [float(np.abs(inter(a) - inter(b)).max() / np.abs(inter(b)).max()) for a, b in zip(res["fp32"], res["fp64"])]
    print("FP32 vs FP64 adjoint step (phi, u, v, w):", rel)
    assert max(rel) <= 1e-4


def test_fp32_gradient_against_fp64():
    """dJ/dA and dJ/dphi0 (along a smooth direction and as fields) of the
    FP32 differentiable advection against the FP64 one: 24^3, 30 steps,
    checkpoints every 7 (uneven last segment)."""
    from fortrancfd_amd.adjoint import advect, l2_loss
    case = _case((24, 24, 24))
    nsteps, period, every = 30, 30 * case.dt, 7
    phi0 = _phi0(case)
    dev = torch.device("cuda")
    with torch.no_grad():
        target = advect(torch.from_numpy(phi0.reshape(-1).copy()).to(dev),
                        torch.tensor(1.05, dtype=torch.float64, device=dev), _solver(case), nsteps, period, nsteps)
    out = {}
    for dt_ in (torch.float64, torch.float32):
        p0 = torch.from_numpy(phi0.reshape(-1).copy()).to(dev, dt_).requires_grad_(True)
        a = torch.tensor(1.0, dtype=dt_, device=dev, requires_grad=True)
        pT = advect(p0, a, _solver(case), nsteps, period, every)
        assert pT.dtype == dt_
        J = l2_loss(pT.double(), target, case)
        J.backward()
        out[dt_] = (float(J), p0.grad.double().cpu().numpy().reshape(case.shape), float(a.grad))
    J64, g64, a64 = out[torch.float64]
    J32, g32, a32 = out[torch.float32]
    v = _smooth(case, np.random.default_rng(3), 1.0)
    d64, d32 = float((g64 * v).sum()), float((g32 * v).sum())
    g = case.ghc
    inter = lambda a: a[g:-g, g:-g, g:-g]
    rel_field = float(np.abs(inter(g32) - inter(g64)).max() / np.abs(inter(g64)).max())
    print("J", J64, J32, "dA", a64, a32, "dphi.v", d64, d32, "field", rel_field)
    assert abs(J32 - J64) <= 1e-4 * abs(J64)
    assert abs(a32 - a64) <= 1e-3 * abs(a64)
    assert abs(d32 - d64) <= 1e-3 * abs(d64)
    assert rel_field <= 1e-2
    m = np.ones(case.shape, dtype=bool)
    m[g:-g, g:-g, g:-g] = False
    assert not g32[m].any()
