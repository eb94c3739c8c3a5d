#!/usr/bin/env bash
# Build the reference RRTMG_SW Fortran (as MiMA compiles it: default REAL is
# 8 bytes, -r8 in CMakeLists.txt:90-97) from the sources where they lie in the
# MiMA reference tree, plus this directory's C-callable driver, into
# oracle/_ref/librrtmg_sw_ref.so.  Test infrastructure only.
#
# Only RRTMG_SW is buildable: its k-table data file (rrtmg_sw_k_g.f90) is in
# the tree.  RRTMG_LW's rrtmg_lw_k_g.f90 is absent, so the LW reference is not
# built.
#
# usage: oracle/build_ref.sh [/root/reference]
set -euo pipefail
REF=${1:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
SW="$REF/src/atmos_param/rrtm_radiation/rrtmg_sw/gcm_model"
FC=${FC:-flang}
if [ ! -d "$SW" ]; then
  echo "build_ref: $SW not found; reference build skipped" >&2
  exit 0
fi
mkdir -p "$OUT/mod" "$OUT/obj"
FFLAGS="-fdefault-real-8 -fPIC -O2 -module-dir $OUT/mod -I$OUT/mod"
MODS="parkind parrrsw rrsw_aer rrsw_cld rrsw_con rrsw_kg16 rrsw_kg17 rrsw_kg18 rrsw_kg19 rrsw_kg20
      rrsw_kg21 rrsw_kg22 rrsw_kg23 rrsw_kg24 rrsw_kg25 rrsw_kg26 rrsw_kg27 rrsw_kg28 rrsw_kg29
      rrsw_ref rrsw_tbl rrsw_vsn rrsw_wvn"
SRCS="rrtmg_sw_setcoef rrtmg_sw_init rrtmg_sw_taumol rrtmg_sw_reftra rrtmg_sw_vrtqdr
      rrtmg_sw_spcvrt rrtmg_sw_cldprop rrtmg_sw_rad.nomcica"
OBJS=""
for m in $MODS; do
  $FC $FFLAGS -c "$SW/modules/$m.f90" -o "$OUT/obj/$m.o"
  OBJS="$OBJS $OUT/obj/$m.o"
done
# the data file is compiled without optimisation, as MiMA does (CMakeLists.txt:70-82)
if [ ! -f "$OUT/obj/rrtmg_sw_k_g.o" ] || [ "$SW/src/rrtmg_sw_k_g.f90" -nt "$OUT/obj/rrtmg_sw_k_g.o" ]; then
  $FC -fdefault-real-8 -fPIC -O0 -module-dir "$OUT/mod" -I"$OUT/mod" -c "$SW/src/rrtmg_sw_k_g.f90" \
      -o "$OUT/obj/rrtmg_sw_k_g.o"
fi
OBJS="$OBJS $OUT/obj/rrtmg_sw_k_g.o"
for s in $SRCS; do
  $FC $FFLAGS -c "$SW/src/$s.f90" -o "$OUT/obj/$s.o"
  OBJS="$OBJS $OUT/obj/$s.o"
done
$FC $FFLAGS -c "$HERE/ref_sw_driver.f90" -o "$OUT/obj/ref_sw_driver.o"
$FC -fdefault-real-8 -fPIC -shared -o "$OUT/librrtmg_sw_ref.so" $OBJS "$OUT/obj/ref_sw_driver.o"
echo "built $OUT/librrtmg_sw_ref.so"
