"""The Fortran binding module (fcfd_c, the ISO_C_BINDING interfaces of
include/fcfd.h, kept as a fortran block of the repository's integration notes)
compiles with AMD flang, links against libfcfd.so and agrees with the C ABI: derived-type sizes
equal the C struct sizes, the field-id parameters equal the header enum, and
a GPU-free entry point (fcfd_prof_nregions) is callable.  No GPU is used."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
LIBDIR = os.path.join(ROOT, "fortrancfd_amd")


def fortran_module() -> str:
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```fortran\n(.*?)```", md, flags=re.S)
    mod = [b for b in blocks if b.lstrip().startswith("module fcfd_c")]
    assert mod, "INTEGRATION.md holds no fcfd_c module block"
    return mod[0]


def test_field_parameters_match_the_header_enum():
    from fortrancfd_amd import types
    for name, val in re.findall(r"FCFD_(\w+)\s*=\s*(\d+)", fortran_module()):
        if name.lower() in types.FIELD_ID:
            assert types.FIELD_ID[name.lower()] == int(val), name


def _fortran_members(typename: str) -> list[str]:
    body = re.search(rf"type, bind\(c\) :: {typename}\n(.*?)end type", fortran_module(), flags=re.S).group(1)
    names = []
    for line in body.splitlines():
        if "::" in line:
            names += [re.sub(r"\(.*\)", "", n).strip() for n in line.split("::")[1].split(",")]
    return names


@pytest.mark.parametrize("fname,pyname", [("fcfd_grid", "FcfdGrid"), ("fcfd_params", "FcfdParams"),
                                          ("fcfd_diag", "FcfdDiag")])
def test_fortran_members_match_the_c_structs(fname, pyname):
    """Same members in the same order as the ctypes mirrors (whose layout
    tests/test_abi.py checks against the header): a size match alone can
    hide a renamed or missing member in the trailing padding."""
    from fortrancfd_amd import types
    assert _fortran_members(fname) == [n for n, _ in getattr(types, pyname)._fields_]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="AMD flang not installed")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libfcfd.so")), reason="libfcfd.so not built")
def test_fortran_binding_compiles_links_and_matches_c():
    from fortrancfd_amd import types
    prog = """
program abi_check
  use fcfd_c
  type(fcfd_grid) :: g
  type(fcfd_params) :: p
  type(fcfd_diag) :: d
  print '(i0)', c_sizeof(g)
  print '(i0)', c_sizeof(p)
  print '(i0)', c_sizeof(d)
  print '(i0)', fcfd_prof_nregions()
end program
"""
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "abi_check.f90")
        open(src, "w").write(fortran_module() + prog)
        exe = os.path.join(t, "abi_check")
        subprocess.run([FLANG, "-module-dir", t, src, "-o", exe, f"-L{LIBDIR}", "-lfcfd", f"-Wl,-rpath,{LIBDIR}"],
                       check=True, capture_output=True, cwd=t)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, cwd=t).stdout.split()
    sizes = [int(v) for v in out[:3]]
    assert sizes == [C.sizeof(types.FcfdGrid), C.sizeof(types.FcfdParams), C.sizeof(types.FcfdDiag)]
    assert int(out[3]) > 10
