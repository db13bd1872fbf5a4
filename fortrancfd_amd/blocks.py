"""Host-side mirror of the reference's multi-block pointer layer
(`DomainDecomposition/3D/lib/pointer/ptr_roots.f90`) over libfcfd.

`Blocks` is the layout of one grid cut into gx x gy x gz blocks (the
reference's grid_x, grid_y, grid_z), each a ghosted array, all of a field's
blocks back to back in one device tensor (a "stack").  Its methods carry the
reference's names: init (the constructor: ptrpart_init with job_init's
ranges), sync (the halo exchange), check_per, plus bc (job_bc(0) on every
block), scatter / gather to and from a whole ghosted field, and
level_set_rk3_solver (the WENO5 SSP-RK3 step block by block with a halo
exchange after each stage).  Tensors are float64 CUDA tensors; every call is
queued on torch's current stream.  No CPU fallback: without the built library
the constructor raises.
"""
from __future__ import annotations

import ctypes as C

from .solver import FcfdError, lib
from .types import Case


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Blocks:
    """gx x gy x gz blocks of `case`'s grid; xper, yper, zper as the
    reference's periodic flags."""

    def __init__(self, case: Case, gx: int, gy: int, gz: int = 1, xper: bool = False, yper: bool = False,
                 zper: bool = False):
        self.case = case
        self.nb = (int(gx), int(gy), int(gz))
        self._grid = case.grid_struct()
        self._params = case.param_struct()
        h = C.c_void_p()
        self._check("fcfd_blocks_create", lib().fcfd_blocks_create(
            C.byref(self._grid), C.byref(self._params), int(gx), int(gy), int(gz), int(bool(xper)),
            int(bool(yper)), int(bool(zper)), C.byref(h)))
        self.h = h
        self.stack_size = int(lib().fcfd_blocks_stack_size(h))

    @staticmethod
    def _check(fn: str, rc: int) -> None:
        if rc != 0:
            raise FcfdError(fn, rc)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().fcfd_blocks_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def range(self, bx: int, by: int, bz: int = 0):
        """((is, ie), (js, je), (ks, ke)) 1-based, and the block's offset in a stack."""
        r = (C.c_int32 * 6)()
        off = C.c_int64()
        self._check("fcfd_blocks_range", lib().fcfd_blocks_range(self.h, bx, by, bz, r, C.byref(off)))
        return ((r[0], r[1]), (r[2], r[3]), (r[4], r[5])), int(off.value)

    def block_view(self, stack, bx: int, by: int, bz: int = 0):
        """block (bx, by, bz) of a stack as a ghosted (k, j, i) view."""
        (i0, i1), (j0, j1), (k0, k1) = self.range(bx, by, bz)[0]
        g = self.case.ghc
        gz = g if self.case.dim == 3 else 0
        ex, ey, ez = i1 - i0 + 1 + 2 * g, j1 - j0 + 1 + 2 * g, k1 - k0 + 1 + 2 * gz
        off = self.range(bx, by, bz)[1]
        return stack[off:off + ex * ey * ez].view(ez, ey, ex)

    def new_stack(self):
        import torch
        return torch.zeros(self.stack_size, dtype=torch.float64, device="cuda")

    def scatter(self, field, stack) -> None:
        self._check("fcfd_blocks_scatter", lib().fcfd_blocks_scatter(self.h, _ptr(field), _ptr(stack), _stream()))

    def gather(self, stack, field) -> None:
        self._check("fcfd_blocks_gather", lib().fcfd_blocks_gather(self.h, _ptr(stack), _ptr(field), _stream()))

    def bc(self, stack) -> None:
        self._check("fcfd_blocks_bc", lib().fcfd_blocks_bc(self.h, _ptr(stack), _stream()))

    def sync(self, stack) -> None:
        self._check("fcfd_blocks_sync", lib().fcfd_blocks_sync(self.h, _ptr(stack), _stream()))

    def check_per(self, stack) -> float:
        e = C.c_double()
        self._check("fcfd_blocks_check_per",
                    lib().fcfd_blocks_check_per(self.h, _ptr(stack), C.byref(e), _stream()))
        return float(e.value)

    def level_set_rk3_solver(self, phi, nu, nv, nw=None) -> None:
        self._check("fcfd_blocks_level_set_rk3_solver", lib().fcfd_blocks_level_set_rk3_solver(
            self.h, _ptr(phi), _ptr(nu), _ptr(nv), _ptr(nw) if nw is not None else None, _stream()))


def block_ranges(n: int, nb: int):
    """job_init's ranges (`DomainDecomposition/3D/lib/data/branches.f90:106-113`):
    block b owns b n / nb + 1 .. (b + 1) n / nb (1-based)."""
    return [(b * n // nb + 1, (b + 1) * n // nb) for b in range(nb)]


__all__ = ["Blocks", "block_ranges"]
