#!/usr/bin/env bash
# Build the reference 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl) with the
# project's test harness main (oracle/ref/harness2d.f90) as an oracle binary under oracle/_ref/.
#
# The reference sources are copied to a scratch directory outside the repository.  Changes to
# the copy: the compile-time PARAMETERs of comdat_param.f90 (grid size, Poisson tolerance and
# sweep cap) and three Intel-dialect constructs flang rejects, rewritten to their standard
# meaning: INT4(x) -> INT(x) (grid.f90, writedata.f90), IF(FLAG) -> IF(FLAG.NE.0)
# (contra_vel.f90), IF(IPOST) -> IF(IPOST.NE.0) (calccoef.f90).  No other source line changes.
#
# usage: build_ref2d.sh LABEL NI NJ NSWPP RESIMAX [HARNESS]   (HARNESS: harness2d (default) or
#        harness2d_post, the trajectory harness with the post-processing phase)
set -euo pipefail
REF_SRC="/root/reference/cylinder/cylinder_v3.0.3_AF_original 4step_face_nl"
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
LABEL=$1; NI=$2; NJ=$3; NSWPP=$4; RESIMAX=$5; HARNESS=${6:-harness2d}
FLANG=${FLANG:-/opt/rocm/lib/llvm/bin/flang}

if [ ! -d "$REF_SRC" ]; then
  echo "reference sources not present; oracle/_ref cannot be built here" >&2
  exit 3
fi
SCRATCH=$(mktemp -d /tmp/fvm_refbuild2d.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
cp "$REF_SRC"/*.f90 "$SCRATCH"/
rm -f "$SCRATCH/Main.f90"
P="$SCRATCH/comdat_param.f90"
sed -E -i \
  -e "s/(PARAMETER ::[[:space:]]*NI[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NI/" \
  -e "s/(PARAMETER ::[[:space:]]*NJ[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NJ/" \
  -e "s/(PARAMETER ::[[:space:]]*NSWPP[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NSWPP/" \
  -e "s/(PARAMETER ::[[:space:]]*RESIMAX[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $RESIMAX/" \
  "$P"
sed -i -e 's/INT4(/INT(/g' "$SCRATCH/grid.f90" "$SCRATCH/writedata.f90"
sed -i -e 's/IF(FLAG) THEN/IF(FLAG.NE.0) THEN/' "$SCRATCH/contra_vel.f90"
sed -i -e 's/IF(IPOST) THEN/IF(IPOST.NE.0) THEN/' "$SCRATCH/calccoef.f90"
grep -E "NI|NJ|NSWPP|RESIMAX" "$P" | grep -v '^[[:space:]]*!' >&2
cp "$HERE/$HARNESS.f90" "$SCRATCH/"
cd "$SCRATCH"
FFLAGS="-O2 -ffp-contract=off"
$FLANG $FFLAGS -c comdat_param.f90 comdat_shared.f90
SRCS=$(ls *.f90 | grep -v -e '^comdat_param.f90$' -e '^comdat_shared.f90$' -e "^$HARNESS.f90\$")
$FLANG $FFLAGS -c $SRCS $HARNESS.f90 2>"$SCRATCH/warn.txt" || { cat "$SCRATCH/warn.txt" >&2; exit 1; }
OUT="$REPO/oracle/_ref/$LABEL"
mkdir -p "$OUT"
$FLANG $FFLAGS -o "$OUT/$HARNESS" *.o
echo "built $OUT/$HARNESS" >&2
