"""CPU checks of the C ABI (no GPU call is made): the built library loads
and exports every function include/fcfd.h declares, and the ctypes mirrors of
the ABI structs in fortrancfd_amd.types have the C layout (sizes and field
offsets measured from the header by gcc)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "fcfd.h")
LIB = os.path.join(ROOT, "fortrancfd_amd", "libfcfd.so")


def declared_functions() -> list[str]:
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"^\s*(?:int|int64_t|const\s+char\s*\*)\s*(fcfd_\w+)\s*\(", src, flags=re.M)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ["fcfd_create", "fcfd_step", "fcfd_level_set_rk3_solver", "fcfd_level_set_rk3_redis",
                 "fcfd_mass_preserving_level_set", "fcfd_ppe_sor_solver", "fcfd_ns_momentum_correction",
                 "fcfd_ppe_solve_rhs", "fcfd_curv", "fcfd_prof_read"]:
        assert must in names
    assert len(names) == len(set(names))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfcfd.so not built (fortrancfd_amd/build.py)")
def test_library_exports_every_declared_function():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def _c_layout(struct: str, fields: list[str]) -> dict:
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "fcfd_types.h")}"',
            'int main(void) {', f'  printf("size %zu\\n", sizeof({struct}));']
    for f in fields:
        prog.append(f'  printf("{f} %zu\\n", offsetof({struct}, {f}));')
    prog += ['  return 0;', '}']
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(c, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-std=c99", c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.parametrize("cname,pyname", [("fcfd_grid", "FcfdGrid"), ("fcfd_params", "FcfdParams"),
                                          ("fcfd_diag", "FcfdDiag"), ("fcfd_case", "FcfdCase")])
def test_ctypes_structs_match_the_c_layout(cname, pyname):
    from fortrancfd_amd import types
    py = getattr(types, pyname)
    fields = [f for f, _ in py._fields_]
    lay = _c_layout(cname, fields)
    assert lay["size"] == C.sizeof(py)
    for f in fields:
        assert lay[f] == getattr(py, f).offset, f


def test_status_and_field_enums_match_the_header():
    from fortrancfd_amd import types
    src = open(os.path.join(ROOT, "include", "fcfd_types.h")).read()
    m = re.search(r"FCFD_NFIELDS", src)
    assert m
    # field ids: the enum body order is the FIELDS list order
    body = src[src.index("enum fcfd_field"):src.index("FCFD_NFIELDS")]
    names = re.findall(r"FCFD_(\w+)", body)
    assert [n.lower() for n in names] == list(types.FIELDS)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfcfd.so not built (fortrancfd_amd/build.py)")
def test_fp16_density_guard():
    """FP16 copies of the density (level 0) and of the level-1 face
    coefficients 2/(rho_a + rho_b) are only used while both fit FP16's
    normal range [6.1e-5, 65504]; host-only, no device call."""
    f = C.CDLL(LIB).fcfd_ppe_fp16_density_ok
    f.argtypes = [C.c_double, C.c_double]
    assert f(1e-3, 1.0) == 1            # the rising bubble, ratio 1000
    assert f(1.0, 1000.0) == 1          # the standalone Poisson density
    assert f(1e-4, 1.0) == 1
    assert f(5e-5, 1.0) == 0            # rho subnormal in FP16
    assert f(1e-5, 1.0) == 0            # 1/rho overflows FP16 on level 1
    assert f(1.0, 1.6e4) == 1 and f(1.0, 2e4) == 0   # 1/rho on level 1 subnormal in FP16
    assert f(0.0, 1.0) == 0 and f(2.0, 1.0) == 0
