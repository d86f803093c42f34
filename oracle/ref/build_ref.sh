#!/bin/bash
# Builds oracle/_ref/adchp_ref: the reference's hot-path Fortran modules,
# compiled unmodified from the sources where they lie under $REF (default
# /root/reference), linked with the harness in this directory.  Outputs go only
# to $OUT (oracle/_ref).  Flags follow the reference's own Release build for
# flang (cmake/architecture.cmake:37-54: -O3 with free line length); no FMA is
# emitted for the generic x86-64 target.  Extra arguments are C objects that
# provide the harness's bind(C) helpers (hash, forcing).
set -euo pipefail
REF=${REF:-/root/reference}
OUT=${OUT:-$(cd "$(dirname "$0")/.." && pwd)/_ref}
HERE=$(cd "$(dirname "$0")" && pwd)
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS="-O3 -ffixed-line-length-none"
[ -d "$REF/src" ] || { echo "reference sources not found at $REF" >&2; exit 2; }
mkdir -p "$OUT/obj"
# dependency order of the modules the hot path USEs (serial build, no MPI/netCDF)
SRCS="src/sizes.F src/constants.F90 src/logging.F90 src/terminate.F90 src/io.F90
thirdparty/KDTREE2/kdtree2.F src/global.F src/boundaries.F src/hashtable.F90 src/mesh.F
src/global_3dvs.F src/itpackv.F src/quadrature.F90 src/internaltide.F90 src/subgridLookup.F
src/nodalattr.F src/vew1d.F90 src/sponge_layer.F90 src/subdomain.F wind/vortex.F
src/owiwind.F src/rs2.F src/owi_ice.F src/nws08.F90 src/wind.F src/momentum.F src/gwce.F
src/wetdry.F"
OBJS=""
for s in $SRCS; do
  o="$OUT/obj/$(basename "$s").o"
  if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
    echo "  FC $s"
    "$FC" -c $FFLAGS -I"$REF/src" -module-dir "$OUT/obj" -I"$OUT/obj" "$REF/$s" -o "$o"
  fi
  OBJS="$OBJS $o"
done
echo "  FC harness.F90"
"$FC" -c -O2 -module-dir "$OUT/obj" -I"$OUT/obj" "$HERE/harness.F90" -o "$OUT/obj/harness.o"
"$FC" -o "$OUT/adchp_ref" "$OUT/obj/harness.o" $OBJS "$@" -lm
echo "  FC jcg_harness.F90"
"$FC" -c -O2 -module-dir "$OUT/obj" -I"$OUT/obj" "$HERE/jcg_harness.F90" -o "$OUT/obj/jcg_harness.o"
"$FC" -o "$OUT/jcg_ref" "$OUT/obj/jcg_harness.o" $OBJS "$@" -lm
echo "built $OUT/adchp_ref $OUT/jcg_ref"
