#!/usr/bin/env bash
# Build the reference RRTMG Fortran (as MiMA compiles it: default REAL is
# 8 bytes, -r8 in CMakeLists.txt:90-97) from the sources where they lie in the
# MiMA reference tree, plus this directory's C-callable drivers, into
# oracle/_ref/librrtmg_sw_ref.so and oracle/_ref/librrtmg_lw_ref.so, and the
# McICA drivers into librrtmg_lw_mcica_ref.so and librrtmg_sw_mcica_ref.so.
# Test infrastructure only.
#
# RRTMG_SW is built whole, its init included: its k-table data file
# (rrtmg_sw_k_g.f90) is in the tree.  For RRTMG_LW the data file
# rrtmg_lw_k_g.f90 is absent, so rrtmg_lw_init.f90 (which calls its
# lw_kgbNN routines) is not built; the LW column code (setcoef, taumol,
# rtrnmr, rtrn, cldprop, rrtmg_lw) is, and ref_lw_driver.f90 fills the module
# state that the init would have set (see that file).
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

LW="$REF/src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model"
mkdir -p "$OUT/mod_lw" "$OUT/obj_lw"
LFLAGS="-fdefault-real-8 -fdefault-double-8 -fPIC -O2 -module-dir $OUT/mod_lw -I$OUT/mod_lw -I$HERE"
LMODS="parkind parrrtm rrlw_cld rrlw_con rrlw_kg01 rrlw_kg02 rrlw_kg03 rrlw_kg04 rrlw_kg05 rrlw_kg06
       rrlw_kg07 rrlw_kg08 rrlw_kg09 rrlw_kg10 rrlw_kg11 rrlw_kg12 rrlw_kg13 rrlw_kg14 rrlw_kg15
       rrlw_kg16 rrlw_ref rrlw_tbl rrlw_vsn rrlw_wvn"
LSRCS="rrtmg_lw_setcoef rrtmg_lw_taumol rrtmg_lw_rtrnmr rrtmg_lw_rtrn rrtmg_lw_cldprop rrtmg_lw_rad.nomcica"
LOBJS=""
for m in $LMODS; do
  $FC $LFLAGS -c "$LW/modules/$m.f90" -o "$OUT/obj_lw/$m.o"
  LOBJS="$LOBJS $OUT/obj_lw/$m.o"
done
for s in $LSRCS; do
  $FC $LFLAGS -c "$LW/src/$s.f90" -o "$OUT/obj_lw/$s.o"
  LOBJS="$LOBJS $OUT/obj_lw/$s.o"
done
$FC $LFLAGS -c "$HERE/ref_lw_driver.f90" -o "$OUT/obj_lw/ref_lw_driver.o"
$FC -fdefault-real-8 -fPIC -shared -Wl,--no-undefined -o "$OUT/librrtmg_lw_ref.so" $LOBJS \
    "$OUT/obj_lw/ref_lw_driver.o" -L"$(dirname "$(command -v $FC)")/../lib" -lflang_rt.runtime 2>/dev/null || \
  $FC -fdefault-real-8 -fPIC -shared -o "$OUT/librrtmg_lw_ref.so" $LOBJS "$OUT/obj_lw/ref_lw_driver.o"
echo "built $OUT/librrtmg_lw_ref.so"

# McICA: its driver rrtmg_lw_rad.f90 defines the module rrtmg_lw_rad as the
# nomcica driver does, so it gets a module directory and library of its own
# (the LW modules, setcoef and taumol objects are shared)
mkdir -p "$OUT/mod_lwm" "$OUT/obj_lwm"
MFLAGS="-fdefault-real-8 -fdefault-double-8 -fPIC -O2 -module-dir $OUT/mod_lwm -I$OUT/mod_lwm -I$OUT/mod_lw -I$HERE"
MOBJS=""
for m in $LMODS; do MOBJS="$MOBJS $OUT/obj_lw/$m.o"; done
MOBJS="$MOBJS $OUT/obj_lw/rrtmg_lw_setcoef.o $OUT/obj_lw/rrtmg_lw_taumol.o"
for s in mcica_random_numbers mcica_subcol_gen_lw rrtmg_lw_cldprmc rrtmg_lw_rtrnmc rrtmg_lw_rad; do
  $FC $MFLAGS -c "$LW/src/$s.f90" -o "$OUT/obj_lwm/$s.o"
  MOBJS="$MOBJS $OUT/obj_lwm/$s.o"
done
$FC $MFLAGS -c "$HERE/ref_lw_mcica_driver.f90" -o "$OUT/obj_lwm/ref_lw_mcica_driver.o"
$FC -fdefault-real-8 -fPIC -shared -o "$OUT/librrtmg_lw_mcica_ref.so" $MOBJS "$OUT/obj_lwm/ref_lw_mcica_driver.o"
echo "built $OUT/librrtmg_lw_mcica_ref.so"

# McICA SW: as for LW, its driver rrtmg_sw_rad.f90 redefines the module
# rrtmg_sw_rad, so it gets its own module directory and library (the SW
# modules, data, init, setcoef, taumol, reftra and vrtqdr objects are shared)
mkdir -p "$OUT/mod_swm" "$OUT/obj_swm"
SMFLAGS="-fdefault-real-8 -fPIC -O2 -module-dir $OUT/mod_swm -I$OUT/mod_swm -I$OUT/mod"
SMOBJS=""
for m in $MODS; do SMOBJS="$SMOBJS $OUT/obj/$m.o"; done
for s in rrtmg_sw_k_g rrtmg_sw_setcoef rrtmg_sw_init rrtmg_sw_taumol rrtmg_sw_reftra rrtmg_sw_vrtqdr; do
  SMOBJS="$SMOBJS $OUT/obj/$s.o"
done
for s in mcica_random_numbers mcica_subcol_gen_sw rrtmg_sw_cldprmc rrtmg_sw_spcvmc rrtmg_sw_rad; do
  $FC $SMFLAGS -c "$SW/src/$s.f90" -o "$OUT/obj_swm/$s.o"
  SMOBJS="$SMOBJS $OUT/obj_swm/$s.o"
done
$FC $SMFLAGS -c "$HERE/ref_sw_mcica_driver.f90" -o "$OUT/obj_swm/ref_sw_mcica_driver.o"
$FC -fdefault-real-8 -fPIC -shared -o "$OUT/librrtmg_sw_mcica_ref.so" $SMOBJS "$OUT/obj_swm/ref_sw_mcica_driver.o"
echo "built $OUT/librrtmg_sw_mcica_ref.so"
