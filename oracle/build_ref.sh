#!/usr/bin/env bash
# Build the compiled Fortran reference harness into oracle/_ref/ (test
# infrastructure only).  The reference sources are compiled where they lie
# under the reference checkout; nothing is copied.  Needs AMD flang
# (/opt/rocm/lib/llvm/bin/flang).  Outputs: oracle/_ref/ref2d, oracle/_ref/ref3d
# (newOpenMP 2D / 3D+IBM through ref_harness.F90) and oracle/_ref/refm3
# (MPLS_RisingBubble/20190612 through m3_harness.F90).
#
# usage: oracle/build_ref.sh [REFERENCE_ROOT]   (default /root/reference)
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ref_root="${1:-/root/reference}"
FC="${FC:-/opt/rocm/lib/llvm/bin/flang}"
out="$here/_ref"
if [ ! -d "$ref_root/Two phase flows/newOpenMP" ]; then
    echo "build_ref: reference checkout not found at $ref_root; skipping" >&2
    exit 0
fi
if [ ! -x "$FC" ]; then
    echo "build_ref: flang not found at $FC; skipping" >&2
    exit 0
fi
mkdir -p "$out"
# reference build flags (N2/makefile:1-6) translated to flang; -O2 without
# -march so no FMA contraction happens on x86-64.
FFLAGS="-O2 -fopenmp"
for d in 2D 3D; do
    if [ "$d" = 2D ]; then src="$ref_root/Two phase flows/newOpenMP/2D"; def=""; exe=ref2d;
    else src="$ref_root/Two phase flows/newOpenMP/3D+IBM"; def="-DHARNESS_3D"; exe=ref3d; fi
    b="$out/build_$d"
    mkdir -p "$b/mod"
    "$FC" $FFLAGS -c -module-dir "$b/mod" -I"$src/lib" "$src/lib/modules.f90" -o "$b/modules.o"
    objs=("$b/modules.o")
    # the vendored VTK writer is output-only and slow to compile: reuse its
    # objects when they are already built
    for f in ir_precision lib_pack_data lib_base64 lib_vtk_io; do
        if [ ! -f "$b/$f.o" ] || [ ! -f "$b/mod/$f.mod" ]; then
            "$FC" -cpp -O1 -c -module-dir "$b/mod" "$src/lib/plot/$f.f90" -o "$b/$f.o"
        fi
        objs+=("$b/$f.o")
    done
    "$FC" $FFLAGS -cpp $def -module-dir "$b/mod" -I"$b/mod" -I"$src/src" "$here/ref_harness.F90" "${objs[@]}" -o "$out/$exe"
    echo "build_ref: built $out/$exe"
done
# the drop-in: the newOpenMP 2D main loop through the fcfd_c binding module
# (the fortran block that starts `module fcfd_c`, extracted below from the
# repository's integration notes) against libfcfd.so
lib="$here/../fortrancfd_amd/libfcfd.so"
if [ -f "$lib" ]; then
    b="$out/build_2D"
    python3 - "$here/../INTEGRATION.md" "$b/fcfd_c.f90" <<'PYEOF'
import re, sys
md = open(sys.argv[1]).read()
blk = [x for x in re.findall(r"```fortran\n(.*?)```", md, flags=re.S) if x.lstrip().startswith("module fcfd_c")][0]
open(sys.argv[2], "w").write(blk)
PYEOF
    "$FC" $FFLAGS -c -module-dir "$b/mod" "$b/fcfd_c.f90" -o "$b/fcfd_c.o"
    "$FC" $FFLAGS -cpp -module-dir "$b/mod" -I"$b/mod" "$here/n2_dropin.F90" "$b/fcfd_c.o" "$b/modules.o" \
        "$b/ir_precision.o" "$b/lib_pack_data.o" "$b/lib_base64.o" "$b/lib_vtk_io.o" \
        -L"$(dirname "$lib")" -lfcfd -Wl,-rpath,"$(cd "$(dirname "$lib")" && pwd)" -o "$out/n2_dropin"
    echo "build_ref: built $out/n2_dropin"
fi
# MPLS_RisingBubble (3D, two bubbles): its modules, then the harness, which
# INCLUDEs the reference's internal subroutines from where they lie
m3="$ref_root/MPLS_RisingBubble/20190612"
if [ -d "$m3" ]; then
    b="$out/build_M3"
    mkdir -p "$b"
    "$FC" $FFLAGS -c -module-dir "$b" "$m3/DATA_MOD.F90" -o "$b/DATA_MOD.o"
    "$FC" $FFLAGS -cpp -module-dir "$b" -I"$b" -I"$m3" "$here/m3_harness.F90" "$b/DATA_MOD.o" -o "$out/refm3" 2> "$b/warnings.txt"
    echo "build_ref: built $out/refm3"
fi
