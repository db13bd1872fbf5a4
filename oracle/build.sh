#!/usr/bin/env bash
# Build the CPU oracle (oracle/_build/liboracle.so) and, when the reference
# checkout and flang are present, the reference harness (oracle/_ref/).
# Test infrastructure only.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
mkdir -p "$here/_build"
gcc -O2 -fPIC -fopenmp -ffp-contract=off -shared -o "$here/_build/liboracle.so.tmp" "$here/fcfd_oracle.c" "$here/m3_oracle.c" -lm
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
echo "oracle: built $here/_build/liboracle.so"
if [ "${FCFD_SKIP_REF:-0}" != "1" ]; then
    "$here/build_ref.sh" "${1:-/root/reference}"
fi
