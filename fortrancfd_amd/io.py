"""On-disk output of the reference's two-phase programs, written from host
copies of the device fields.

* `write_vts`: the reference's `plot` (`Two phase flows/newOpenMP/2D/src/
  others/plot.f90:1-52`, 3D `3D+IBM/src/others/plot.f90:1-57`): a VTK XML
  StructuredGrid in Lib_VTK_IO's "raw" appended encoding (`.../lib/plot/
  lib_vtk_io.f90` VTK_INI_XML :909-1037, VTK_GEO_XML_STRG :1348-1504,
  VTK_DAT_XML :2641-2724, VTK_VAR_XML :2726-2774 and :3335-3392,
  VTK_END_XML :4721-4869).  The domain is cut into num_of_plot pieces per
  axis, neighbouring pieces sharing one node layer; nodes are the cell
  centres x_i = xstart + (i - 1/2) dx (`lib/data/tree.f90:146-161`); point
  data Phi, Pressure and the cell-centred velocity (nvel).  The 3D+IBM
  build also writes the immersed-boundary "Solid" field after Phi; this
  path carries no immersed boundary, so it is written only when given.
* `read_vts`: reads such a file back (header offsets + appended blocks).
* `MVLossLog`: the `_MVloss.plt` history (`lib/data/tree.f90:120-122`,
  `src/others/output.f90:8`): time, percent loss of mass and of volume
  against the values after `problem_init`.
* `write_tecplot_point`: the MPLS_RisingBubble Tecplot ASCII snapshot
  (`MPLS_RisingBubble/20190612/PLOT.F90:26-63`): one POINT zone over the
  cells and their first ghost layer, 4-significant-digit E fields.
"""
from __future__ import annotations

import os
import re
import struct
from typing import Dict, Iterable, List, Sequence, Tuple

import numpy as np

from .types import Case

_NL = b"\n"


def _pieces(n: int, num_of_plot: int) -> List[Tuple[int, int]]:
    """1-based node ranges of the pieces along one axis: piece q covers
    q*n/np + 1 .. (q+1)*n/np, lowered by one node for q > 0 so that
    neighbouring pieces share their boundary layer (plot.f90:20-28)."""
    out = []
    for q in range(num_of_plot):
        a = q * n // num_of_plot + 1
        b = (q + 1) * n // num_of_plot
        if q > 0:
            a -= 1
        out.append((a, b))
    return out


def _ext(*r) -> str:
    """Lib_VTK_IO writes extents through str(n=...), which prefixes '+'
    to non-negative integers (ir_precision.f90 str_I4P)."""
    return " ".join(("+%d" % v) if v >= 0 else ("%d" % v) for v in r)


def write_vts(path: str, case: Case, phi: np.ndarray, p: np.ndarray, vel: Sequence[np.ndarray],
              num_of_plot: int = 1, solid: np.ndarray = None) -> None:
    """Write one plot snapshot.  phi, p and the cell-centred velocity
    components are ghosted fields of `case` (C order, as Solver.pull
    returns them); only the interior nodes are written.  solid: an
    immersed-boundary field written after Phi as "Solid", the 3D+IBM
    build's layout (this path has none; pass zeros to reproduce that
    layout exactly)."""
    dim = case.dim
    nx, ny = case.nx, case.ny
    nz = case.nz if dim == 3 else 1
    inter = lambda a: np.asarray(case.interior(np.asarray(a, dtype=np.float64)))
    ph, pr = inter(phi), inter(p)
    vx, vy = inter(vel[0]), inter(vel[1])
    vz = inter(vel[2]) if dim == 3 else np.zeros_like(vx)
    so = inter(solid) if solid is not None else None
    if dim == 2:   # treat 2D fields as one k plane: arrays indexed [k, j, i]
        ph, pr, vx, vy, vz = (a[None, :, :] for a in (ph, pr, vx, vy, vz))
        so = so[None, :, :] if so is not None else None
    scalars = [("Phi", ph)] + ([("Solid", so)] if so is not None else []) + [("Pressure", pr)]
    xs = case.xstart + (np.arange(1, nx + 1) - 0.5) * case.dx
    ys = case.ystart + (np.arange(1, ny + 1) - 0.5) * case.dy
    zs = case.zstart + (np.arange(1, nz + 1) - 0.5) * case.dz if dim == 3 else np.zeros(1)

    head: List[bytes] = []
    blocks: List[bytes] = []
    offset = 0

    def add_block(a: np.ndarray) -> int:
        nonlocal offset
        raw = np.ascontiguousarray(a, dtype="<f8").tobytes()
        blocks.append(struct.pack("<i", len(raw)) + raw)
        off = offset
        offset += 4 + len(raw)
        return off

    def line(indent: int, text: str) -> None:
        head.append((" " * indent + text).encode() + _NL)

    head.append(b'<?xml version="1.0"?>' + _NL)
    line(0, '<VTKFile type="StructuredGrid" version="0.1" byte_order="LittleEndian">')
    line(2, '<StructuredGrid WholeExtent="%s">' % _ext(1, nx, 1, ny, 1, nz))
    zr = _pieces(nz, num_of_plot) if dim == 3 else [(1, 1)]
    yr_all = _pieces(ny, num_of_plot)
    xr_all = _pieces(nx, num_of_plot)
    # piece loops: x outermost, then y, then z (plot.f90 loop order)
    for (x1, x2) in xr_all:
        for (y1, y2) in yr_all:
            for (z1, z2) in zr:
                sl = (slice(z1 - 1, z2), slice(y1 - 1, y2), slice(x1 - 1, x2))
                Z, Y, X = np.meshgrid(zs[z1 - 1:z2], ys[y1 - 1:y2], xs[x1 - 1:x2], indexing="ij")
                # Fortran order of a piece: i fastest, then j, then k
                pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
                line(4, '<Piece Extent="%s">' % _ext(x1, x2, y1, y2, z1, z2))
                line(6, "<Points>")
                line(8, '<DataArray type="Float64" NumberOfComponents="3" Name="Points" '
                        'format="appended" offset="%d"/>' % add_block(pts))
                line(6, "</Points>")
                line(6, "<PointData>")
                for name, a in scalars:
                    line(8, '<DataArray type="Float64" Name="%s" NumberOfComponents="1" '
                            'format="appended" offset="%d"/>' % (name, add_block(a[sl].ravel())))
                v = np.stack([vx[sl].ravel(), vy[sl].ravel(), vz[sl].ravel()], axis=1)
                line(8, '<DataArray type="Float64" Name="Velocity" NumberOfComponents="3" '
                        'format="appended" offset="%d"/>' % add_block(v))
                line(6, "</PointData>")
                line(4, "</Piece>")
    line(2, "</StructuredGrid>")
    line(2, '<AppendedData encoding="raw">')
    with open(path, "wb") as fh:
        fh.write(b"".join(head))
        fh.write(b"_")
        fh.write(b"".join(blocks))
        fh.write(_NL)
        fh.write(b"  </AppendedData>" + _NL)
        fh.write(b"</VTKFile>" + _NL)


def read_vts(path: str) -> Dict[str, object]:
    """Parse a raw-appended StructuredGrid file: returns the whole extent and,
    per piece, its extent and every data array (Points as (N, 3))."""
    data = open(path, "rb").read()
    mark = data.index(b'<AppendedData encoding="raw">')
    start = data.index(b"_", mark) + 1
    header = data[:mark].decode()

    def block(off: int, ncomp: int) -> np.ndarray:
        (nb,) = struct.unpack_from("<i", data, start + off)
        a = np.frombuffer(data, dtype="<f8", count=nb // 8, offset=start + off + 4)
        return a.reshape(-1, ncomp) if ncomp > 1 else a.copy()

    whole = tuple(int(v) for v in re.search(r'WholeExtent="([^"]+)"', header).group(1).split())
    pieces = []
    for m in re.finditer(r'<Piece Extent="([^"]+)">(.*?)</Piece>', header, re.S):
        ext = tuple(int(v) for v in m.group(1).split())
        arrays = {}
        for d in re.finditer(r'<DataArray type="Float64"([^/]*)/>', m.group(2)):
            attrs = dict(re.findall(r'(\w+)="([^"]*)"', d.group(1)))
            arrays[attrs["Name"]] = block(int(attrs["offset"]), int(attrs["NumberOfComponents"]))
        pieces.append({"extent": ext, "arrays": arrays})
    return {"whole_extent": whole, "pieces": pieces}


def _fortran_e(v: float, digits: int, width: int) -> str:
    """Fortran Ew.d output (0.dddE+xx form, right-aligned in w columns)."""
    if v == 0.0:
        mant, ex = 0.0, 0
    else:
        ex = int(np.floor(np.log10(abs(v)))) + 1
        mant = v / 10.0 ** ex
        if round(abs(mant), digits) >= 1.0:   # rounding carried into the next decade
            mant /= 10.0
            ex += 1
    s = ("%." + str(digits) + "f") % abs(mant)
    s = s[1:] if s.startswith("0") else s   # ".dddd"
    txt = ("-" if v < 0 else "") + "0" + s + ("E%+03d" % ex)
    return txt.rjust(width)


def list_directed(v: float) -> str:
    """A real(8) as Fortran list-directed output of the reference build
    (flang) writes it: the shortest digits that round-trip, fixed notation
    without the leading zero for 0.1 <= |v| < 1e15 (".125", "3.",
    "123456."), else d.ddE+xx with at least two exponent digits ("1.E-05",
    "2.5E-12")."""
    if v == 0.0:
        return "0."
    sign = "-" if v < 0 else ""
    mant, _, ex = ("%r" % float(abs(v))).partition("e")
    if "." not in mant:
        mant += ".0"
    ip, fp = mant.split(".")
    digits = (ip + fp).lstrip("0").rstrip("0") or "0"
    # decimal exponent: value = 0.d1d2... x 10^e10
    lead = len(ip.lstrip("0")) if ip.strip("0") else -(len(fp) - len(fp.lstrip("0")))
    e10 = lead + (int(ex) if ex else 0)
    a = abs(v)
    if 0.1 <= a < 1.0e15:
        if e10 <= 0:
            body = "." + "0" * (-e10) + digits
        elif e10 >= len(digits):
            body = digits + "0" * (e10 - len(digits)) + "."
        else:
            body = digits[:e10] + "." + digits[e10:]
        return sign + body
    x = e10 - 1
    return sign + digits[0] + "." + digits[1:] + ("E%s%02d" % ("+" if x >= 0 else "-", abs(x)))


class MVLossLog:
    """The `_MVloss.plt` history file of the reference (tree.f90:120-122:
    a Tecplot header line; output.f90:8: one row per call with time,
    100 (imass - mass)/imass and 100 (ivol - vol)/ivol).  imass and ivol
    are the values after problem_init (problem_init.f90:81-84)."""

    def __init__(self, path: str, imass: float, ivol: float):
        self.path = path
        self.imass, self.ivol = float(imass), float(ivol)
        with open(path, "w") as fh:
            fh.write(' variables = "T" "Loss of mass" "Loss of Volume" \n')

    def row(self, time: float, mass: float, vol: float) -> Tuple[float, float]:
        lm = 100.0 * (self.imass - mass) / self.imass
        lv = 100.0 * (self.ivol - vol) / self.ivol
        with open(self.path, "a") as fh:
            fh.write(" " + " ".join(list_directed(v) for v in (time, lm, lv)) + "\n")
        return lm, lv

    @staticmethod
    def read(path: str) -> np.ndarray:
        rows = [list(map(float, ln.split())) for ln in open(path).read().splitlines()[1:] if ln.strip()]
        return np.asarray(rows, dtype=np.float64).reshape(-1, 3)


def write_tecplot_point(path: str, case: Case, title: str, fields: Dict[str, np.ndarray]) -> None:
    """One POINT zone over cells 0 .. n+1 per axis (the interior and its
    first ghost layer) with X, Y, Z and the given ghosted fields, each value
    as Fortran E14.4 (PLOT.F90:26-63; its IBM_SOLVER = 0 branch writes
    "X" "Y" "Z" followed by the level-set fields).  Coordinates of the ghost
    layer follow the cell-centre spacing."""
    if case.dim != 3:
        raise ValueError("the Tecplot snapshot is the 3D MPLS_RisingBubble writer")
    g = case.ghc
    X, Y, Z = case.coords()
    sl = (slice(g - 1, case.nz + g + 1), slice(g - 1, case.ny + g + 1), slice(g - 1, case.nx + g + 1))
    cols = [X[sl], Y[sl], Z[sl]] + [np.asarray(f, dtype=np.float64).reshape(case.shape)[sl] for f in fields.values()]
    flat = np.stack([c.ravel() for c in cols], axis=1)
    names = ["X", "Y", "Z"] + list(fields.keys())
    with open(path, "w") as fh:
        fh.write(' TITLE = "T=%s" \n' % title)
        fh.write(" VARIABLES =  " + " ".join('"%s"' % n for n in names) + "\n")
        fh.write('ZONE T= "NUMERICAL" , I=%4d, J=%4d, K=%4d, F= POINT\n' % (case.nx + 2, case.ny + 2, case.nz + 2))
        for r in flat:
            fh.write("".join(_fortran_e(v, 4, 14) for v in r) + "\n")
