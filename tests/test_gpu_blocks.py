"""The multi-block layout on the device (fcfd_blocks_* through the C ABI)
against the restatement of the reference's ptr_roots.f90 / job_bc
(oracle/blocks.py, pinned in test_blocks_oracle.py), and the block-by-block
WENO5 SSP-RK3 step against the whole-field step.

* ranges and stack offsets: job_init's ranges, the stack layout;
* job_bc(0) and the halo exchange on seeded blocks (ghosts and all random):
  bit for bit the oracle's job_bc and its settled sync (three passes of the
  reference's loop; a fourth changes nothing); face ghosts also equal one
  pass;
* check_per: the oracle's sum to 1e-12 relative (another summation order);
* scatter / gather: bit for bit;
* level_set_rk3_solver block by block with the halo exchange after each
  stage: bit for bit fcfd_level_set_rk3_solver on the whole field (itself
  within 1e-12 of the oracle, test_gpu_parity.py), over three steps, on
  even and uneven decompositions; on periodic layouts the result does not
  depend on the decomposition.
"""
import numpy as np
import pytest

from cases import bubble, case2d, case3d, case3d_odd, node_velocities
from fortrancfd_amd.types import bc_copy
from oracle import blocks as ob

pytestmark = pytest.mark.gpu

LAYOUTS = [
    ("2d_2x3", lambda: case2d(32), (2, 3, 1), (False, False, False)),
    ("2d_3x2_xper", lambda: case2d(21), (3, 2, 1), (True, False, False)),
    ("3d_2x2x3", lambda: case3d(16), (2, 2, 3), (False, False, False)),
    ("3d_odd_3x2x4", case3d_odd, (3, 2, 4), (False, False, False)),
    ("3d_per_xyz", lambda: case3d(12), (2, 3, 2), (True, True, True)),
    ("3d_2x1x1_zper", lambda: case3d(12), (2, 1, 1), (False, False, True)),
]


def _blocks(case, nb, per):
    from fortrancfd_amd.blocks import Blocks
    return Blocks(case, *nb, xper=per[0], yper=per[1], zper=per[2])


def _layout(case, nb, per):
    return ob.Layout((case.nx, case.ny, case.nz if case.dim == 3 else 1), nb, dim=case.dim, per=per)


def _to_stack(B, lay, blocks):
    import torch
    st = torch.zeros(B.stack_size, dtype=torch.float64)
    for b in lay.keys():
        off = B.range(*b)[1]
        a = blocks[b].reshape(-1)
        st[off:off + a.size] = torch.from_numpy(a)
    return st.cuda()


def _from_stack(B, lay, st):
    h = st.cpu().numpy()
    out = {}
    for b in lay.keys():
        off = B.range(*b)[1]
        shp = lay.shape(b)
        out[b] = h[off:off + int(np.prod(shp))].reshape(shp)
    return out


@pytest.mark.parametrize("tag,mk,nb,per", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_ranges_and_stack_size(tag, mk, nb, per):
    case = mk()
    B, lay = _blocks(case, nb, per), _layout(case, nb, per)
    total = 0
    for b in lay.keys():
        (i, j, k), off = B.range(*b)
        assert (i, j, k) == (lay.r[0][b[0]], lay.r[1][b[1]], lay.r[2][b[2]])
        total += int(np.prod(lay.shape(b)))
    assert B.stack_size == total


@pytest.mark.parametrize("tag,mk,nb,per", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_bc_sync_check_per_match_the_oracle(tag, mk, nb, per):
    import torch
    case = mk()
    B, lay = _blocks(case, nb, per), _layout(case, nb, per)
    rng = np.random.default_rng(11)
    blocks = {b: rng.standard_normal(lay.shape(b)) for b in lay.keys()}
    # job_bc
    st = _to_stack(B, lay, blocks)
    B.bc(st)
    torch.cuda.synchronize()
    want = {b: a.copy() for b, a in blocks.items()}
    ob.job_bc(lay, want)
    got = _from_stack(B, lay, st)
    for b in lay.keys():
        assert np.array_equal(got[b], want[b]), ("bc", b)
    # the halo exchange from random ghosts (physical ghosts are copied as they are)
    st = _to_stack(B, lay, blocks)
    B.sync(st)
    torch.cuda.synchronize()
    got = _from_stack(B, lay, st)
    one = {b: a.copy() for b, a in blocks.items()}
    ob.sync(lay, one)
    settled = {b: a.copy() for b, a in one.items()}
    ob.sync(lay, settled)
    ob.sync(lay, settled)
    again = {b: a.copy() for b, a in settled.items()}
    ob.sync(lay, again)
    for b in lay.keys():
        assert np.array_equal(again[b], settled[b]), ("oracle not settled", b)
        assert np.array_equal(got[b], settled[b]), ("sync", b)
        ez, ey, ex = lay.shape(b)
        kk, jj, ii = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
        out = ((ii < lay.g) | (ii >= ex - lay.g)).astype(int) + ((jj < lay.g) | (jj >= ey - lay.g)).astype(int)
        if lay.gz:
            out += ((kk < lay.gz) | (kk >= ez - lay.gz)).astype(int)
        assert np.array_equal(got[b][out <= 1], one[b][out <= 1]), ("face ghosts after one pass", b)
    # check_per on the random state and on the synchronised one
    st = _to_stack(B, lay, blocks)
    e_dev, e_ref = B.check_per(st), ob.check_per(lay, blocks)
    assert abs(e_dev - e_ref) <= 1e-12 * max(e_ref, 1e-300), (e_dev, e_ref)
    if any(per):
        assert e_ref > 0.0
    B.sync(st)
    assert B.check_per(st) == 0.0


@pytest.mark.parametrize("tag,mk,nb,per", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_scatter_gather(tag, mk, nb, per):
    import torch
    case = mk()
    B, lay = _blocks(case, nb, per), _layout(case, nb, per)
    F = np.random.default_rng(3).standard_normal(case.shape)
    st = B.new_stack()
    B.scatter(torch.from_numpy(F).cuda(), st)
    torch.cuda.synchronize()
    got, want = _from_stack(B, lay, st), ob.from_global(lay, F)
    for b in lay.keys():
        assert np.array_equal(got[b], want[b]), b
    G = torch.zeros(case.shape, dtype=torch.float64, device="cuda")
    B.gather(st, G)
    g, gz = case.ghc, (case.ghc if case.dim == 3 else 0)
    Gi = G.cpu().numpy()
    if case.dim == 3:
        assert np.array_equal(Gi[gz:-gz, g:-g, g:-g], F[gz:-gz, g:-g, g:-g])
    else:
        assert np.array_equal(Gi[g:-g, g:-g], F[g:-g, g:-g])


def _advect_blocks(case, nb, per, phi, vel, steps):
    import torch
    B = _blocks(case, nb, per)
    st = {}
    for k, a in [("phi", phi)] + list(vel.items()):
        st[k] = B.new_stack()
        B.scatter(torch.from_numpy(np.ascontiguousarray(a)).cuda(), st[k])
        B.sync(st[k])
    for _ in range(steps):
        B.level_set_rk3_solver(st["phi"], st["nu"], st["nv"], st.get("nw"))
    out = torch.zeros(case.shape, dtype=torch.float64, device="cuda")
    B.gather(st["phi"], out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("tag,mk,nb,per", [x for x in LAYOUTS if not any(x[3])],
                         ids=[x[0] for x in LAYOUTS if not any(x[3])])
def test_blockwise_advection_equals_the_whole_field(tag, mk, nb, per):
    from fortrancfd_amd import Solver
    case = mk()
    phi = bc_copy(case, bubble(case, 0.3, 1))
    vel = {k: bc_copy(case, v) for k, v in node_velocities(case, 3).items()}
    s = Solver(case)
    s.push("phi", phi)
    for k, v in vel.items():
        s.push(k, v)
    for _ in range(3):
        s.level_set_rk3_solver()
    whole = s.pull("phi")
    got = _advect_blocks(case, nb, per, phi, vel, 3)
    I = case.interior
    assert np.array_equal(I(got), I(whole)), float(np.abs(I(got) - I(whole)).max())
    assert not np.array_equal(I(got), I(phi))


@pytest.mark.parametrize("tag,mk,nb,per", [x for x in LAYOUTS if any(x[3])], ids=[x[0] for x in LAYOUTS if any(x[3])])
def test_periodic_advection_does_not_depend_on_the_decomposition(tag, mk, nb, per):
    case = mk()
    phi = bubble(case, 0.3, 1)
    vel = node_velocities(case, 4)
    one = _advect_blocks(case, (1, 1, 1), per, phi, vel, 3)
    got = _advect_blocks(case, nb, per, phi, vel, 3)
    I = case.interior
    assert np.array_equal(I(got), I(one)), float(np.abs(I(got) - I(one)).max())
