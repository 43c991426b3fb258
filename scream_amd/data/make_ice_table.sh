#!/bin/bash
# Regenerates p3_lookup_table_1.dat-v4.1.1 with the reference's own generator
# (components/eam/tools/create_p3_lookupTable/create_p3_lookupTable_1.f90-v4.1),
# compiled as its header asks (double precision), then prepends the
# "VERSION 4.1.1" header that both table readers require
# (components/eamxx/src/physics/p3/impl/p3_table_ice_impl.hpp:37-41,
#  components/eam/src/physics/p3/scream/micro_p3.F90:147-157).
# About 16 minutes on one core.  Needs /root/reference and amdflang.
set -euo pipefail
REF=${REF:-/root/reference}
FLANG=${FLANG:-/opt/rocm/llvm/bin/amdflang}
HERE=$(cd "$(dirname "$0")" && pwd)
WORK=$(mktemp -d)
trap 'rm -rf "$WORK"' EXIT
ln -s "$REF/components/eam/tools/create_p3_lookupTable/create_p3_lookupTable_1.f90-v4.1" "$WORK/gen.f90"
"$FLANG" -O2 -fdefault-real-8 -fdefault-double-8 "$WORK/gen.f90" -o "$WORK/gen"
(cd "$WORK" && ./gen > gen.log)
{ echo "VERSION 4.1.1"; cat "$WORK/p3_lookup_table_1.dat-v4.1"; } > "$HERE/p3_lookup_table_1.dat-v4.1.1"
sha256sum "$HERE/p3_lookup_table_1.dat-v4.1.1"
