#!/bin/sh
# Test infrastructure only.
#
# Compiles the UNMODIFIED reference AMUN Fortran modules straight from their
# sources in the reference tree (no copies are made) together with this
# project's own harness program (oracle/ref_harness.F90), for NDIMS=2 and
# NDIMS=3, with amdflang.  Outputs go to oracle/_ref/ only:
#   oracle/_ref/amun_ref_2d.x, oracle/_ref/amun_ref_3d.x
#
# The reference's own make system is bypassed on purpose: its makefile adds
# -D${COMPILER}, and -DGNU makes parameters.F90:118-122 declare
# "intrinsic iargc, getarg", which flang rejects.  No -DMPI, no -DHDF5 (I/O
# becomes a no-op).  -O2 without -march keeps the x86-64 baseline ISA (no FMA
# contraction), so the binaries also run on other hosts.
#
# domains.F90 is compiled at -O0: amdflang 22 at -O2 drops the low-side
# periodic neighbour links that setup_domain_default wires with
# cshift(...,-1) (domains.F90:336-360), so the lowest blocks would keep
# zero-gradient ghosts.  That module only runs once at start-up.
#
# usage: oracle/build_ref.sh [reference_root]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${1:-${AMUN_REFERENCE_ROOT:-/root/reference}}
SRC="$REF/src"
OUT="$HERE/_ref"
FC=${FC:-amdflang}
if [ ! -f "$SRC/evolution.F90" ]; then
  echo "build_ref.sh: reference sources not found under $SRC" >&2
  exit 2
fi
MODS="constants algebra timers blocks mpitools parameters coordinates \
equations user_problem gravity interpolations boundaries domains shapes \
operators random problems refinement mesh schemes sources evolution io \
integrals"
mkdir -p "$OUT"
for N in 2 3; do
  OBJ="$OUT/obj${N}d"
  mkdir -p "$OBJ"
  OBJS=""
  for m in $MODS; do
    OPT=-O2
    [ "$m" = domains ] && OPT=-O0
    $FC -cpp $OPT -DNDIMS=$N -module-dir "$OBJ" -c "$SRC/$m.F90" -o "$OBJ/$m.o"
    OBJS="$OBJS $OBJ/$m.o"
  done
  $FC -cpp -O2 -DNDIMS=$N -module-dir "$OBJ" -I"$OBJ" \
      -c "$HERE/ref_harness.F90" -o "$OBJ/ref_harness.o"
  $FC -O2 -o "$OUT/amun_ref_${N}d.x" "$OBJ/ref_harness.o" $OBJS
done
echo "built $OUT/amun_ref_2d.x $OUT/amun_ref_3d.x"
