"""On-disk formats (fortrancfd_amd.io) against the reference's own writers.

Where the compiled reference harness is available (oracle/_ref), the plot
snapshot must be byte-identical to the one the reference's `plot` writes
through Lib_VTK_IO, and the `_MVloss.plt` header and rows identical to its
`manager_init` header and `output` rows (flang list-directed reals).  The
closed-form checks (round trip, piece split, offsets, Fortran E format) run
without it.
"""
import re

import numpy as np
import pytest

from fortrancfd_amd.io import (MVLossLog, _fortran_e, _pieces, list_directed, read_vts, write_tecplot_point,
                               write_vts)
from fortrancfd_amd.types import Case, bc_copy, bubble_phi, read_reference_input
from oracle import refrun

VTS_CASES = [(2, 10, 14, 1, 3), (3, 8, 6, 10, 2)]


def _case(dim, nx, ny, nz, npl):
    c = Case(dim=dim, nx=nx, ny=ny, nz=nz, xend=1.0, yend=ny / nx if dim == 2 else 0.75, zend=1.25,
             name="cmp")
    c.num_of_plot = npl
    return c


def _fields(case, seed):
    rng = np.random.default_rng(seed)
    names = ["phi", "p", "nu", "nv"] + (["nw"] if case.dim == 3 else [])
    return {k: rng.standard_normal(case.size).reshape(case.shape) for k in names}


def _vel(case, f):
    return [f["nu"], f["nv"]] + ([f["nw"]] if case.dim == 3 else [])


@pytest.mark.parametrize("dim,nx,ny,nz,npl", VTS_CASES)
def test_vts_identical_to_reference_plot(tmp_path, dim, nx, ny, nz, npl):
    if not refrun.available(dim):
        pytest.skip("reference harness not built (oracle/build_ref.sh)")
    case = _case(dim, nx, ny, nz, npl)
    f = _fields(case, 11 + dim)
    ref = refrun.run(case, "plot", f)["_files"]["cmp_000.vts"]
    path = tmp_path / "mine.vts"
    # the 3D+IBM build writes its immersed-boundary field (zero here) after Phi
    write_vts(str(path), case, f["phi"], f["p"], _vel(case, f), npl,
              solid=np.zeros(case.shape) if dim == 3 else None)
    assert path.read_bytes() == ref


def test_piece_split_closed_form():
    # plot.f90:20-28: q*n/np + 1 .. (q+1)*n/np, lower end minus one for q > 0
    assert _pieces(10, 3) == [(1, 3), (3, 6), (6, 10)]
    assert _pieces(8, 2) == [(1, 4), (4, 8)]
    assert _pieces(64, 1) == [(1, 64)]


@pytest.mark.parametrize("dim,nx,ny,nz,npl", VTS_CASES)
def test_vts_round_trip(tmp_path, dim, nx, ny, nz, npl):
    case = _case(dim, nx, ny, nz, npl)
    f = _fields(case, 5)
    path = str(tmp_path / "a.vts")
    write_vts(path, case, f["phi"], f["p"], _vel(case, f), npl)
    d = read_vts(path)
    nzz = nz if dim == 3 else 1
    assert d["whole_extent"] == (1, nx, 1, ny, 1, nzz)
    ph = case.interior(f["phi"]).reshape(nzz, ny, nx)
    pr = case.interior(f["p"]).reshape(nzz, ny, nx)
    vx = case.interior(f["nu"]).reshape(nzz, ny, nx)
    xs = (np.arange(1, nx + 1) - 0.5) * case.dx
    assert len(d["pieces"]) == npl ** dim
    for pc in d["pieces"]:
        x1, x2, y1, y2, z1, z2 = pc["extent"]
        sl = (slice(z1 - 1, z2), slice(y1 - 1, y2), slice(x1 - 1, x2))
        a = pc["arrays"]
        assert np.array_equal(a["Phi"], ph[sl].ravel())
        assert np.array_equal(a["Pressure"], pr[sl].ravel())
        assert np.array_equal(a["Velocity"][:, 0], vx[sl].ravel())
        if dim == 2:
            assert not a["Velocity"][:, 2].any() and not a["Points"][:, 2].any()
        # nodes are the cell centres, i fastest
        n = x2 - x1 + 1
        assert np.allclose(a["Points"][:n, 0], xs[x1 - 1:x2], rtol=0, atol=1e-15)


def test_vts_offsets_follow_raw_encoding(tmp_path):
    """Lib_VTK_IO raw appended data: each block is an Int32 byte count and
    the data; a DataArray's offset is the running sum of 4 + bytes
    (byte_update, lib_vtk_io.f90:566-596)."""
    case = _case(3, 4, 4, 4, 1)
    f = _fields(case, 2)
    path = str(tmp_path / "b.vts")
    write_vts(path, case, f["phi"], f["p"], _vel(case, f), 1)
    txt = open(path, "rb").read()
    head = txt[: txt.index(b'<AppendedData encoding="raw">')].decode()
    offs = [int(v) for v in re.findall(r'offset="(\d+)"', head)]
    n = 64
    assert offs == [0, 4 + 24 * n, 8 + 32 * n, 12 + 40 * n]   # Points, Phi, Pressure, Velocity
    assert head.splitlines()[2] == '  <StructuredGrid WholeExtent="+1 +4 +1 +4 +1 +4">'


def test_list_directed_against_reference(tmp_path):
    """Rows of _MVloss.plt from the reference's `output` vs MVLossLog, and
    the flang list-directed form of the time column over its regimes."""
    if not refrun.available(2):
        pytest.skip("reference harness not built (oracle/build_ref.sh)")
    case = Case(dim=2, nx=8, ny=16, nz=1, xend=1.0, yend=2.0, name="cmp")
    phi = bc_copy(case, bubble_phi(case, 0.25, (0.5, 0.5)))
    for t in [0.125, 1e-05, 0.01, 0.1, 3.0, 12.5, 1234567.0, 1e15, 2.5e-12, 0.0]:
        out = refrun.run(case, "output", {"phi": phi}, imass=0.9, time=t)
        lines = out["_files"]["cmp_MVloss.plt"].decode().splitlines(keepends=True)
        assert lines[0] == ' variables = "T" "Loss of mass" "Loss of Volume" \n'
        log_path = str(tmp_path / "probe_MVloss.plt")
        log = MVLossLog(log_path, 0.9, 1.01 * out["_vol"])
        log.row(t, out["_mass"], out["_vol"])
        mine = open(log_path).read().splitlines(keepends=True)
        assert mine[0] == lines[0]
        assert mine[1] == lines[1], (t, mine[1], lines[1])


def test_list_directed_closed_form():
    cases = {1e-05: "1.E-05", 0.1: ".1", 1.0: "1.", 123456.0: "123456.", 1e20: "1.E+20", -0.5: "-.5",
             0.0: "0.", 2.5e-12: "2.5E-12", 0.125: ".125"}
    for v, s in cases.items():
        assert list_directed(v) == s
        assert float(s) == v


def test_mvloss_values(tmp_path):
    path = str(tmp_path / "x_MVloss.plt")
    log = MVLossLog(path, imass=2.0, ivol=4.0)
    log.row(0.5, 1.9, 4.2)
    log.row(1.0, 2.0, 4.0)
    rows = MVLossLog.read(path)
    assert np.allclose(rows, [[0.5, 5.0, -5.0], [1.0, 0.0, 0.0]], rtol=1e-15, atol=1e-13)


def test_tecplot_point(tmp_path):
    case = Case(dim=3, nx=3, ny=2, nz=2, xend=1.0, yend=2.0 / 3.0, zend=2.0 / 3.0)
    phi = np.full(case.shape, 0.5)
    path = str(tmp_path / "MC_3D_000.dat")
    write_tecplot_point(path, case, "000", {"phi": phi})
    lines = open(path).read().splitlines()
    assert lines[0] == ' TITLE = "T=000" '
    assert lines[1] == ' VARIABLES =  "X" "Y" "Z" "phi"'
    assert lines[2] == 'ZONE T= "NUMERICAL" , I=   5, J=   4, K=   4, F= POINT'
    assert len(lines) == 3 + 5 * 4 * 4
    # first point: the ghost cell (0,0,0) at x = -dx/2; Fortran E14.4
    assert lines[3] == _fortran_e(-case.dx / 2, 4, 14) + _fortran_e(-case.dy / 2, 4, 14) + \
        _fortran_e(-case.dz / 2, 4, 14) + "    0.5000E+00"
    assert _fortran_e(0.5, 4, 14) == "    0.5000E+00"
    assert _fortran_e(-123.456, 4, 14) == "   -0.1235E+03"
    assert _fortran_e(0.99996, 4, 14) == "    0.1000E+01"


def test_reference_input_carries_plot_schedule(tmp_path):
    c = Case(dim=2, nx=64, ny=128, nz=1, xend=1.0, yend=2.0, name="rb")
    c.num_of_plot, c.t2s, c.t2p = 4, 3.0, 0.25
    p = tmp_path / "in.txt"
    p.write_text(c.to_reference_input())
    r = read_reference_input(str(p), 2)
    assert (r.num_of_plot, r.t2s, r.t2p) == (4, 3.0, 0.25)
