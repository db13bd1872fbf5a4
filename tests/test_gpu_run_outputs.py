"""The reference time loop with its outputs (fortrancfd_amd.run) on the
device: plot snapshots every t2p, the legacy mesh header and one
_MVloss.plt row per step, read back and checked against the solver state."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_run_writes_reference_outputs(tmp_path):
    from fortrancfd_amd.io import MVLossLog, read_vts
    from fortrancfd_amd.run import run
    from fortrancfd_amd.types import Case, bc_copy, bubble_phi
    case = Case(dim=2, nx=32, ny=64, nz=1, xend=1.0, yend=2.0, name="rb", ppe_tol=1e-10)
    # plot.f90:10 writes when |time - pid*t2p| <= dt: 0, 2dt, 4dt here
    case.num_of_plot, case.t2p = 2, 2.4 * case.dt
    case.t2s = 4.5 * case.dt
    phi0 = bc_copy(case, bubble_phi(case, 0.25, (0.5, 0.5)))
    d = run(case, phi0, str(tmp_path))
    assert d["iter"] == 5                                  # stops once time > t2s
    rows = MVLossLog.read(str(tmp_path / "rb_MVloss.plt"))
    assert rows.shape == (5, 3)
    assert np.allclose(rows[:, 0], case.dt * np.arange(1, 6), rtol=1e-14)
    assert np.all(np.abs(rows[:, 1:]) < 1.0)               # percent losses stay small
    snaps = sorted(p.name for p in tmp_path.glob("rb_*.vts"))
    assert snaps == ["rb_000.vts", "rb_001.vts", "rb_002.vts"]   # t = 0, 2dt, 4dt
    v = read_vts(str(tmp_path / "rb_002.vts"))
    assert v["whole_extent"] == (1, 32, 1, 64, 1, 1) and len(v["pieces"]) == 4
    assert all(np.isfinite(pc["arrays"]["Phi"]).all() for pc in v["pieces"])
    assert (tmp_path / "rb_mesh.vtk").read_text().startswith("# vtk DataFile Version 3.0")
