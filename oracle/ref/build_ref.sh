#!/usr/bin/env bash
# Builds the reference SHARK Fortran step (read-only sources under /root/reference/src)
# with AMD flang into oracle/_ref/<name>/, linked to the dump harness ref_harness.f90
# (a test driver that calls the reference's own read_params / setup / solve and writes
# the full u_prim state after a chosen number of steps).
#
# The object list and the source search order follow bin/Makefile:35,67-85 of the
# reference (VPATH order; dust objects only when NDUST>0).  gfortran-only flags of the
# reference Makefile (bin/Makefile:47) are replaced by their flang equivalents; -fopenmp
# is always given because commons.f90:12 uses OMP_LIB unconditionally.  Runs must be
# single-threaded (OMP_NUM_THREADS=1): the reference has two OpenMP data races
# (godunov.f90:43 dx_loc, dust_growth.f90:63 m2).
#
# usage: build_ref.sh SETUP NX NY NDUST GEOM GRIDSPACE SOLVER SOLVERDUST [extra -D flags]
set -euo pipefail
REF=${SHARK_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
SETUP=$1; NX=$2; NY=$3; NDUST=$4; GEOM=$5; GRIDSPACE=$6; SOLVER=$7; SOLVERDUST=$8
shift 8
EXTRA="$*"
NAME=${REF_NAME:-${SETUP}_${NX}x${NY}_n${NDUST}_g${GEOM}${GRIDSPACE}_s${SOLVER}${SOLVERDUST}}
OUT=$HERE/../_ref/$NAME
mkdir -p "$OUT/obj"
if [ ! -d "$REF/src" ]; then echo "reference sources not found at $REF" >&2; exit 2; fi

# GRAVITY defaults to 0; an extra -DGRAVITY=... replaces it (flang keeps the first definition)
GRAV="-DGRAVITY=0"
case " $EXTRA " in *" -DGRAVITY="*) GRAV="" ;; esac
DEFINES="-DNDUST=$NDUST -DNX=$NX -DNY=$NY -DGRIDSPACE=$GRIDSPACE -DGEOM=$GEOM -DMHD=0 -DOPENMP=1 \
 -DSOLVER=$SOLVER -DSOLVERDUST=$SOLVERDUST -DSOLVERB=1 -DNGHOST=2 -DTURB=0 $GRAV -DNDUSTPSCAL=0 $EXTRA"
FC=${FC:-flang}
FFLAGS="-cpp -O3 -fopenmp $DEFINES -module-dir $OUT/obj"

VPATH="$REF/src $REF/src/commons $REF/src/grid $REF/src/turb $REF/src/solver $REF/src/setups/$SETUP \
 $REF/src/gravity $REF/src/hydro $REF/src/chemistry $REF/src/dust"
MODOBJ="precision phys_const hydro_commons setup_commons gravity_commons ionisation_commons random newton_raphson"
[ "$NDUST" != "0" ] && MODOBJ="$MODOBJ dust_commons"
MODOBJ="$MODOBJ commons parameters units boundary_types slope_limiters allocate_init ionisation_utils functions_NR \
 ionisation grid solvers eos setup courant godunov boundaries write_output read_output source_terms force_kicks fargo_scheme"
[ "$NDUST" != "0" ] && MODOBJ="$MODOBJ dust_init dust_drag dust_growth"
MODOBJ="$MODOBJ gravity_utils solve read_params"

find_src() {
  for d in $VPATH; do
    if [ -f "$d/$1.f90" ]; then echo "$d/$1.f90"; return 0; fi
  done
  echo "missing source $1.f90" >&2; return 1
}

OBJS=""
# REF_FARGO_FLAGS: extra flags for fargo_scheme.f90 alone, e.g. "-Done=1.0d0 -fdefault-real-8" to
# give its undeclared `one` (implicit REAL(4), never assigned, :57,69) the value 1 and its
# undeclared `eps` (:47) double precision (-fdefault-double-8 keeps 1.0d0 literals at kind 8)
for o in $MODOBJ; do
  src=$(find_src "$o")
  OFLAGS=""
  [ "$o" = "fargo_scheme" ] && OFLAGS="${REF_FARGO_FLAGS:-}"
  $FC $FFLAGS $OFLAGS -c "$src" -o "$OUT/obj/$o.o"
  OBJS="$OBJS $OUT/obj/$o.o"
done
HFLAGS=""
[ "$SETUP" = "disk_polar" ] && HFLAGS="-DREF_DISK"
[ "$SETUP" = "SI" ] && HFLAGS="-DREF_SI"
[ "$SETUP" = "SOD" ] && HFLAGS="-DREF_SOD"
[ "$SETUP" = "KH" ] && HFLAGS="-DREF_KH"
[ "$SETUP" = "disk" ] && HFLAGS="-DREF_CDISK"
[ "$SETUP" = "Collapse_1D" ] && HFLAGS="-DREF_COLLAPSE"
HARNESS=${REF_HARNESS:-ref_harness.f90}
$FC $FFLAGS $HFLAGS -c "$HERE/$HARNESS" -o "$OUT/obj/ref_harness.o"
$FC -fopenmp $OBJS "$OUT/obj/ref_harness.o" -o "$OUT/shark_ref"
echo "$OUT/shark_ref"
