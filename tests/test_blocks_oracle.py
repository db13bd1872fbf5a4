"""The multi-block oracle (oracle/blocks.py, a restatement of the
reference's ptr_roots.f90 / job_init / job_bc) pinned by what the layout
fixes mathematically.  The reference's multi-block code cannot be built here
(its tree module needs LAPACK), so these are its pins:

* job_init's ranges tile 1..N with lengths that differ by at most one;
* job_bc(0) makes every ghost the block's interior cell at the clamped index;
* after job_bc on every block and the halo exchange, the blocks hold the
  restriction of ONE global field: the global field's interior, continued by
  the periodic wrap on periodic axes and by the zero-gradient copy on the
  others.  One pass of the reference's sync gets every face ghost (one index
  outside the block) right; the settled state (three passes) every ghost;
  a further pass changes nothing;
* check_per is 0 on the settled state and sums the squared perturbation of a
  periodic ghost.
"""
import itertools

import numpy as np
import pytest

from oracle import blocks as ob

LAYOUTS = [
    ("2d_2x3", (32, 40, 1), (2, 3, 1), 2, (False, False, False)),
    ("2d_3x2_per", (21, 17, 1), (3, 2, 1), 2, (True, False, False)),
    ("3d_2x2x3", (16, 16, 24), (2, 2, 3), 3, (False, False, False)),
    ("3d_3x2x4_uneven", (20, 18, 37), (3, 2, 4), 3, (False, False, False)),
    ("3d_per_xyz", (12, 15, 18), (2, 3, 2), 3, (True, True, True)),
    ("3d_one_block_per_z", (12, 12, 12), (2, 1, 1), 3, (False, False, True)),
]


def _global(n, dim, per, seed):
    """a seeded global ghosted field whose ghosts are the periodic wrap on
    periodic axes and the zero-gradient copy on the others (applied axis by
    axis over the full extents, as the reference's bc does)"""
    g, gz = 3, (3 if dim == 3 else 0)
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((n[2] + 2 * gz, n[1] + 2 * g, n[0] + 2 * g))
    for axis, ga in ((0, g), (1, g), (2, gz)):
        if ga == 0:
            continue
        ax = 2 - axis
        m = n[axis]
        for q in range(ga):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], hi[ax] = q, ga + m + q
            src_lo = [slice(None)] * 3
            src_hi = [slice(None)] * 3
            if per[axis]:
                src_lo[ax], src_hi[ax] = m + q, ga + q
            else:
                src_lo[ax], src_hi[ax] = ga, ga + m - 1
            F[tuple(lo)] = F[tuple(src_lo)]
            F[tuple(hi)] = F[tuple(src_hi)]
    return F


@pytest.mark.parametrize("n", [1, 7, 16, 37, 100])
@pytest.mark.parametrize("nb", [1, 2, 3, 5])
def test_ranges_tile_the_axis(n, nb):
    r = ob.ranges(n, nb)
    assert r[0][0] == 1 and r[-1][1] == n
    for (a, b), (c, d) in zip(r, r[1:]):
        assert c == b + 1
    lens = [b - a + 1 for a, b in r]
    assert max(lens) - min(lens) <= 1 and sum(lens) == n


@pytest.mark.parametrize("tag,n,nb,dim,per", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_bc_is_the_clamped_interior(tag, n, nb, dim, per):
    lay = ob.Layout(n, nb, dim=dim, per=per)
    rng = np.random.default_rng(1)
    B = {b: rng.standard_normal(lay.shape(b)) for b in lay.keys()}
    ref = {b: a.copy() for b, a in B.items()}
    ob.job_bc(lay, B)
    for b in lay.keys():
        ez, ey, ex = lay.shape(b)
        g, gz = lay.g, lay.gz
        k = np.clip(np.arange(ez), gz, ez - 1 - gz)
        j = np.clip(np.arange(ey), g, ey - 1 - g)
        i = np.clip(np.arange(ex), g, ex - 1 - g)
        assert np.array_equal(B[b], ref[b][np.ix_(k, j, i)])


@pytest.mark.parametrize("tag,n,nb,dim,per", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_sync_restricts_one_global_field(tag, n, nb, dim, per):
    lay = ob.Layout(n, nb, dim=dim, per=per)
    F = _global(n, dim, per, 2)
    want = ob.from_global(lay, F)
    # interiors from F, ghosts garbage, then bc on every block and the sync
    B = {}
    for b, a in want.items():
        x = np.full_like(a, np.nan)
        s = (slice(lay.gz, a.shape[0] - lay.gz), slice(lay.g, a.shape[1] - lay.g), slice(lay.g, a.shape[2] - lay.g))
        x[s] = a[s]
        B[b] = x
    ob.job_bc(lay, B)
    ob.sync(lay, B)
    # after one pass: every cell with at most one index outside its block
    for b in lay.keys():
        ez, ey, ex = lay.shape(b)
        kk, jj, ii = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
        out = ((ii < lay.g) | (ii >= ex - lay.g)).astype(int) + ((jj < lay.g) | (jj >= ey - lay.g)).astype(int)
        if lay.gz:
            out += ((kk < lay.gz) | (kk >= ez - lay.gz)).astype(int)
        face = out <= 1
        assert np.array_equal(B[b][face], want[b][face]), b
    for _ in range(2):
        ob.sync(lay, B)
    for b in lay.keys():
        assert np.array_equal(B[b], want[b]), b
    ob.sync(lay, B)
    for b in lay.keys():
        assert np.array_equal(B[b], want[b]), b
    assert ob.check_per(lay, B) == 0.0


def test_check_per_sums_the_squared_perturbation():
    lay = ob.Layout((12, 15, 18), (2, 3, 2), dim=3, per=(True, False, False))
    F = _global(lay.n, 3, lay.per, 5)
    B = ob.from_global(lay, F)
    assert ob.check_per(lay, B) == 0.0
    b = (0, 1, 1)
    B[b][4, 5, 1] += 0.5     # an x-low ghost of the first block (local i = 1 < g)
    assert ob.check_per(lay, B) == 0.25
    lay0 = ob.Layout((12, 15, 18), (2, 3, 2), dim=3, per=(False, False, False))
    assert ob.check_per(lay0, B) == 0.0


def test_single_block_sync_is_the_periodic_wrap_only():
    """one block per axis: the sync only wraps the periodic axes"""
    n = (10, 9, 8)
    for per in itertools.product((False, True), repeat=3):
        lay = ob.Layout(n, (1, 1, 1), dim=3, per=per)
        F = _global(n, 3, (False, False, False), 7)   # zero-gradient ghosts everywhere
        B = ob.from_global(lay, F)
        ob.sync(lay, B)
        for _ in range(2):
            ob.sync(lay, B)
        assert np.array_equal(B[(0, 0, 0)], _global(n, 3, per, 7)), per
