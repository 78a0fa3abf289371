#!/bin/bash
# Compiles the reference's own Fortran sources for the oracle's golden vectors.
# (HOP's chain, its spectrum and sampled-table modules, and SRC/ray_mod.f90 with
# solvers/rt/radau_mod.f90: the files that compile without the DISPATCH framework.)
#
# The HOP chain (experiments/OpenMULTI3D/HOP of the reference) is compiled from
# its sources where they lie, with amdflang, -DHOP_STANDALONE (HOP's own
# standalone configuration, HOP/standalone/Makefile) and -O2 without fast-math
# (the reference's production flag policy, config/compiler/gfortran/hoppe.mkf).
# Then the harness programs under oracle/ref_harness (this repository's code)
# are linked against those objects.  Outputs go to oracle/_ref/ only.
# Test infrastructure: nothing in the product links these objects.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF=${REF:-/root/reference/experiments/OpenMULTI3D}
H="$REF/HOP"
OUT="$HERE/_ref"
FC=${FC:-amdflang}
FFLAGS="-cpp -DHOP_STANDALONE -O2 -fPIC -fopenmp -w"
if [ ! -d "$H" ]; then
  echo "build_ref.sh: $H not found; nothing to build" >&2
  exit 0
fi
command -v "$FC" >/dev/null || { echo "build_ref.sh: $FC not found" >&2; exit 0; }
mkdir -p "$OUT/mod"
# HOP standalone modules, dependency order
MODS="console_mod namelist_check_mod qlg_mod hop_env_mod const_mod op_math_mod interpol_mod
handyfuncs_mod linelist_reader_mod anstee_mod hvcs_mod level_mod line_types_mod bbtrans_mod
ion_mod element_mod molecule_mod hprofl_mod linelist_mod"
stamp="$OUT/.stamp"
for m in $MODS; do
  src="$H/$m.f90"
  obj="$OUT/$m.o"
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
    $FC $FFLAGS -I"$H" -module-dir "$OUT/mod" -c "$src" -o "$obj"
  fi
done
# the spectrum and sampled-table modules (HOP), the angular quadratures
# (SRC/ray_mod.f90 with solvers/rt/radau_mod.f90): each compiles on its own
EXTRA="bftrans_mod trans_mod spectrum_mod optab_mod"
for m in $EXTRA; do
  src="$H/$m.f90"
  obj="$OUT/$m.o"
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
    $FC $FFLAGS -I"$H" -I"$OUT/mod" -module-dir "$OUT/mod" -c "$src" -o "$obj"
  fi
done
RADAU="${REF%/experiments/OpenMULTI3D}/solvers/rt/radau_mod.f90"
for pair in "radau_mod:$RADAU" "ray_mod:$REF/SRC/ray_mod.f90"; do
  m=${pair%%:*}
  src=${pair#*:}
  obj="$OUT/$m.o"
  if [ -f "$src" ] && { [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; }; then
    $FC $FFLAGS -I"$H" -I"$OUT/mod" -module-dir "$OUT/mod" -c "$src" -o "$obj"
  fi
done
# harness programs: hop_golden links the HOP chain, rt_golden adds the modules above
for hsrc in "$HERE"/ref_harness/*.f90; do
  [ -f "$hsrc" ] || continue
  name=$(basename "$hsrc" .f90)
  exe="$OUT/$name.x"
  objs=$(for m in $MODS; do echo "$OUT/$m.o"; done)
  if [ "$name" = rt_golden ]; then
    [ -f "$OUT/ray_mod.o" ] || continue
    objs="$objs $(for m in $EXTRA radau_mod ray_mod; do echo "$OUT/$m.o"; done)"
  fi
  if [ ! -f "$exe" ] || [ "$hsrc" -nt "$exe" ] || [ "$OUT/linelist_mod.o" -nt "$exe" ]; then
    $FC $FFLAGS -I"$H" -I"$OUT/mod" -module-dir "$OUT/mod" "$hsrc" $objs -o "$exe"
  fi
done
touch "$stamp"
