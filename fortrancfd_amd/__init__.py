"""fortrancfd_amd — MI355X-native two-phase Navier-Stokes step (level set +
MPLS + 1/rho multigrid pressure solve + projection) of the FortranCFD
reference, behind a C ABI (include/fcfd.h) implemented in HIP for gfx950."""
from .types import FIELDS, PPE_FP64, PPE_MIXED, Case, bubble_phi, read_reference_input, rising_bubble  # noqa: F401
from .solver import FcfdError, Solver, lib  # noqa: F401

__all__ = ["Case", "Solver", "FcfdError", "FIELDS", "PPE_FP64", "PPE_MIXED", "lib", "bubble_phi",
           "read_reference_input", "rising_bubble"]
