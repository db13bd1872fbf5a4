"""fcfd_ls_mv_noalloc, the entry point that allocates nothing, against
fcfd_ls_mv of an ordinary context (and of single contexts per case for a
batched 2D stack): the property fields H, delta, sign, rho, mu and the mass
and volume sums must be bit for bit those of the existing entry point.  Two
calls in a row, the second on a smaller input, so the second call runs on
the scratch array the first one left behind.  Called through the C ABI with
ctypes; inputs and outputs are torch tensors on the device."""
import ctypes as C

import numpy as np
import pytest

from cases import bubble, case2d, case3d

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

OUTS = ("heavy", "delta", "sign", "rho", "mu")


def _noalloc(cases, phis):
    """one fcfd_ls_mv_noalloc call on the stacked level sets of `cases`
    (one 3D case, or one or more 2D cases of one grid)."""
    from fortrancfd_amd.solver import lib
    from fortrancfd_amd.types import FcfdCase
    L = lib()
    fn = L.fcfd_ls_mv_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p] + [C.c_void_p] * 7 + [C.c_void_p]
    fn.restype = C.c_int
    c0 = cases[0]
    n = len(cases)
    dev = torch.device("cuda")
    phi = torch.from_numpy(np.ascontiguousarray(np.stack(phis) if n > 1 else phis[0])).to(dev)
    outs = {k: torch.full_like(phi, np.nan) for k in OUTS}
    sums = torch.full((2 * n,), np.nan, dtype=torch.float64, device=dev)
    grid, prm = c0.grid_struct(), c0.param_struct()
    arr = None
    if n > 1:
        arr = (FcfdCase * n)()
        for q, c in enumerate(cases):
            arr[q].rho_12, arr[q].mu_12 = c.rho_12, c.mu_12
            arr[q].re, arr[q].we, arr[q].fr = c.re, c.we, c.fr
    torch.cuda.synchronize()
    rc = fn(C.byref(grid), C.byref(prm), n, C.cast(arr, C.c_void_p) if arr is not None else None,
            phi.data_ptr(), *(outs[k].data_ptr() for k in OUTS), sums.data_ptr(), None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in outs.items()}, sums.cpu().numpy()


def _reference(case, phi):
    """fcfd_ls_mv of an ordinary context."""
    from fortrancfd_amd import Solver
    s = Solver(case)
    s.push("phi", phi)
    m, v = s.ls_mv()
    out = {k: s.pull(k) for k in OUTS}
    s.close()
    return out, (m, v)


def _check(cases, phis):
    got, sums = _noalloc(cases, phis)
    for q, (c, phi) in enumerate(zip(cases, phis)):
        ref, (m, v) = _reference(c, phi)
        for k in OUTS:
            a = got[k][q] if len(cases) > 1 else got[k]
            assert np.array_equal(a, ref[k]), (k, q)
        assert sums[2 * q] == m and sums[2 * q + 1] == v, (q, sums[2 * q:2 * q + 2], m, v)


def test_noalloc_3d_two_calls_second_smaller():
    big, small = case3d(32), case3d(12)
    _check([big], [bubble(big, 0.3, 1)])
    _check([small], [bubble(small, 0.2, 2)])


def test_noalloc_batched_2d_two_calls_second_smaller():
    first = [case2d(40, rho_12=r, mu_12=m, name="b") for r, m in ((1e-3, 1e-2), (0.05, 0.2), (0.3, 0.5))]
    second = [case2d(24, rho_12=r, mu_12=m, name="b") for r, m in ((0.02, 0.1), (1e-3, 1e-2))]
    _check(first, [bubble(c, 0.1 * (q + 1), q) for q, c in enumerate(first)])
    _check(second, [bubble(c, 0.2, 5 + q) for q, c in enumerate(second)])


def test_noalloc_rejects_more_cases_than_its_scratch():
    from fortrancfd_amd.solver import lib
    from fortrancfd_amd.types import FcfdCase
    L = lib()
    c = case2d(8)
    n = 65
    arr = (FcfdCase * n)()
    dev = torch.device("cuda")
    phi = torch.zeros(n * c.size, dtype=torch.float64, device=dev)
    sums = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    fn = L.fcfd_ls_mv_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p] + [C.c_void_p] * 7 + [C.c_void_p]
    rc = fn(C.byref(c.grid_struct()), C.byref(c.param_struct()), n, C.cast(arr, C.c_void_p), phi.data_ptr(),
            None, None, None, None, None, sums.data_ptr(), None)
    assert rc == 1   # FCFD_ERR_ARG, before any launch
