"""Restatement of the reference's multi-block layer in numpy: the block
ranges of job_init, the zero-gradient ghost copy job_bc(0), the halo exchange
ptrpart_sync and ptrpart_check_periodic
(`DomainDecomposition/3D/lib/data/branches.f90:84-121,240-290`,
`DomainDecomposition/3D/lib/pointer/ptr_roots.f90:157-404`).

TEST INFRASTRUCTURE ONLY: the checker the HIP layout (libfcfd fcfd_blocks_*)
is compared against.  Only tests/ import it.

The reference's multi-block code is not built here: its tree module calls
LAPACK's dgesv (`.../lib/data/tree_multigrid.f90:159,227`), which this image
does not have.  The restatement keeps the reference's loop order: blocks in
(idz, idy, idx) order, in each the +x, -x, +y, -y, +z, -z ghost slabs over
the block's full extents of the other axes, each copied from the
neighbour's array, then the periodic wraps per axis.  Slab copies are
vectorised; within one slab the source is another block's array, so the
element order inside it does not matter.

Blocks are dicts {(bx, by, bz): array}, each array of shape
(ez, ey, ex) in (k, j, i) order covering the global 1-based indices
is - g .. ie + g (z: ks - gz .. ke + gz; gz = 0 in 2D).
"""
from __future__ import annotations

import numpy as np


def ranges(n: int, nb: int):
    """job_init (branches.f90:106-113): IS = ID N / G + 1, IE = (ID + 1) N / G."""
    return [(b * n // nb + 1, (b + 1) * n // nb) for b in range(nb)]


class Layout:
    def __init__(self, n, nb, g: int = 3, dim: int = 3, per=(False, False, False)):
        self.n, self.nb, self.g, self.dim = tuple(n), tuple(nb), g, dim
        self.gz = g if dim == 3 else 0
        self.per = tuple(bool(p) for p in per)
        self.r = [ranges(n[a], nb[a]) for a in range(3)]

    def keys(self):
        for bz in range(self.nb[2]):
            for by in range(self.nb[1]):
                for bx in range(self.nb[0]):
                    yield bx, by, bz

    def shape(self, b):
        (i0, i1), (j0, j1), (k0, k1) = self.r[0][b[0]], self.r[1][b[1]], self.r[2][b[2]]
        return (k1 - k0 + 1 + 2 * self.gz, j1 - j0 + 1 + 2 * self.g, i1 - i0 + 1 + 2 * self.g)

    def lo(self, b):
        """global 1-based index of element (0, 0, 0) of block b, (i, j, k)"""
        return self.r[0][b[0]][0] - self.g, self.r[1][b[1]][0] - self.g, self.r[2][b[2]][0] - self.gz

    def new(self, fill=0.0):
        return {b: np.full(self.shape(b), fill) for b in self.keys()}

    # a cell of block b by global 1-based indices
    def at(self, blocks, b, i, j, k):
        li, lj, lk = self.lo(b)
        return blocks[b][k - lk, j - lj, i - li]


def _sl(lay, b, axis, g0, g1):
    """index tuple (k, j, i) of block b selecting global indices g0..g1 along
    axis (0 = i, 1 = j, 2 = k) and everything along the other two"""
    lo = lay.lo(b)
    s = [slice(None)] * 3
    s[2 - axis] = slice(g0 - lo[axis], g1 - lo[axis] + 1)
    return tuple(s)


def job_bc(lay: Layout, blocks: dict) -> None:
    """job_bc(0) on every block: x, then y, then z zero-gradient copies over
    the full extents (branches.f90:246-281)."""
    for b in lay.keys():
        d = blocks[b]
        for axis in range(3):
            ga = lay.gz if axis == 2 else lay.g
            s, e = lay.r[axis][b[axis]]
            for q in range(1, ga + 1):
                d[_sl(lay, b, axis, s - q, s - q)] = d[_sl(lay, b, axis, s, s)]
                d[_sl(lay, b, axis, e + q, e + q)] = d[_sl(lay, b, axis, e, e)]


def sync(lay: Layout, blocks: dict) -> None:
    """ptrpart_sync (ptr_roots.f90:157-320): one pass in the reference's order."""
    nb = lay.nb
    for b in lay.keys():
        d = blocks[b]
        for axis in range(3):
            ga = lay.gz if axis == 2 else lay.g
            if ga == 0:
                continue
            s, e = lay.r[axis][b[axis]]
            if b[axis] < nb[axis] - 1:     # +axis ghosts from the next block
                o = list(b); o[axis] += 1; o = tuple(o)
                d[_sl(lay, b, axis, e + 1, e + ga)] = blocks[o][_sl(lay, o, axis, e + 1, e + ga)]
            if b[axis] > 0:                # -axis ghosts from the previous block
                o = list(b); o[axis] -= 1; o = tuple(o)
                d[_sl(lay, b, axis, s - ga, s - 1)] = blocks[o][_sl(lay, o, axis, s - ga, s - 1)]
    # periodic wraps (ptr_roots.f90:255-316), axis by axis
    for axis in range(3):
        if not lay.per[axis]:
            continue
        ga = lay.gz if axis == 2 else lay.g
        last = nb[axis] - 1
        eL = lay.r[axis][last][1]
        for b in lay.keys():
            if b[axis] != 0:
                continue
            L = list(b); L[axis] = last; L = tuple(L)
            for q in range(1, ga + 1):
                blocks[b][_sl(lay, b, axis, 1 - q, 1 - q)] = blocks[L][_sl(lay, L, axis, eL + 1 - q, eL + 1 - q)]
                blocks[L][_sl(lay, L, axis, eL + q, eL + q)] = blocks[b][_sl(lay, b, axis, q, q)]


def check_per(lay: Layout, blocks: dict) -> float:
    """ptrpart_check_periodic (ptr_roots.f90:322-404): the squared
    differences of the periodic ghost layers and the cells they mirror, summed
    over the periodic axes."""
    err = 0.0
    for axis in range(3):
        if not lay.per[axis]:
            continue
        ga = lay.gz if axis == 2 else lay.g
        last = lay.nb[axis] - 1
        eL = lay.r[axis][last][1]
        for b in lay.keys():
            if b[axis] != 0:
                continue
            L = list(b); L[axis] = last; L = tuple(L)
            for q in range(1, ga + 1):
                a1 = blocks[b][_sl(lay, b, axis, 1 - q, 1 - q)] - blocks[L][_sl(lay, L, axis, eL + 1 - q, eL + 1 - q)]
                a2 = blocks[L][_sl(lay, L, axis, eL + q, eL + q)] - blocks[b][_sl(lay, b, axis, q, q)]
                err += float((a1 ** 2).sum() + (a2 ** 2).sum())
    return err


def from_global(lay: Layout, field) -> dict:
    """every cell of every block from a ghosted global array (k, j, i) with
    ghost width g (gz in z) at the same global index"""
    field = np.asarray(field)
    if field.ndim == 2:    # a 2D field: one plane, no z ghosts
        field = field[None]
    out = {}
    for b in lay.keys():
        li, lj, lk = lay.lo(b)
        ez, ey, ex = lay.shape(b)
        # global 1-based index m sits at array position m - 1 + g
        out[b] = np.array(field[lk - 1 + lay.gz:lk - 1 + lay.gz + ez, lj - 1 + lay.g:lj - 1 + lay.g + ey,
                                li - 1 + lay.g:li - 1 + lay.g + ex])
    return out
