"""Host-side checks of the batched 2D context wrapper (no GPU call): the
cases of one batch must share the grid and every shared parameter, and the
per-case parameter struct carries what may differ."""
import pytest

from cases import case2d


def test_batch_rejects_mismatched_cases():
    from fortrancfd_amd.solver import BatchSolver
    a = case2d(16)
    with pytest.raises(ValueError):
        BatchSolver([a, case2d(16, cfl=0.2)])        # a different time step
    with pytest.raises(ValueError):
        BatchSolver([a, case2d(32)])                 # a different grid
    with pytest.raises(ValueError):
        BatchSolver([a, case2d(16, redis_interval=3)])
    with pytest.raises(ValueError):
        BatchSolver([])


def test_case_struct_carries_the_per_case_parameters():
    from fortrancfd_amd.solver import BatchSolver
    c = case2d(16, rho_12=0.02, mu_12=0.03, re=50.0, we=7.0, fr=2.0)
    s = BatchSolver._case_struct(c)
    assert (s.rho_12, s.mu_12, s.re, s.we, s.fr) == (0.02, 0.03, 50.0, 7.0, 2.0)
