#!/usr/bin/env bash
# Build the reference 3D LES solver (LES/v2.3_LES-Reynolds stresses) with the project's test
# harness main (oracle/ref/harness3d.f90) as an oracle binary under oracle/_ref/.
#
# The reference sources are copied to a scratch directory outside the repository; only the
# compile-time PARAMETER constants of comdat_param.f90 are substituted (grid size, a single
# MPI rank, Poisson tolerance/sweep cap, restart on, Re, dt, reference area).  No source line
# of the solver itself is changed.  The binary (never committed) is written to
# oracle/_ref/<label>/harness3d.
#
# usage: build_ref.sh LABEL NI NJ NK RE DTIME NSWPP RESIMAX USE_LES REF_A
set -euo pipefail
REF_SRC="/root/reference/LES/v2.3_LES-Reynolds stresses"
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
LABEL=$1; NI=$2; NJ=$3; NK=$4; RE=$5; DTIME=$6; NSWPP=$7; RESIMAX=$8; USE_LES=$9; REF_A=${10}
FLANG=${FLANG:-/opt/rocm/lib/llvm/bin/flang}
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}

if [ ! -d "$REF_SRC" ]; then
  echo "reference sources not present; oracle/_ref cannot be built here" >&2
  exit 3
fi
SCRATCH=$(mktemp -d /tmp/fvm_refbuild.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
cp "$REF_SRC"/*.f90 "$SCRATCH"/
rm -f "$SCRATCH/main.f90"
P="$SCRATCH/comdat_param.f90"
sed -E -i \
  -e "s/(PARAMETER ::[[:space:]]*NI[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NI/" \
  -e "s/(PARAMETER ::[[:space:]]*NJ[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NJ/" \
  -e "s/(PARAMETER ::[[:space:]]*NK[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NK/" \
  -e "s/(PARAMETER ::[[:space:]]*NDIMJ[[:space:]]*=)[[:space:]]*[0-9]+/\1 1/" \
  -e "s/(PARAMETER ::[[:space:]]*NDIMK[[:space:]]*=)[[:space:]]*[0-9]+/\1 1/" \
  -e "s/(PARAMETER ::[[:space:]]*NSWPP[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NSWPP/" \
  -e "s/(PARAMETER ::[[:space:]]*IRESTART[[:space:]]*=)[[:space:]]*[0-9]+/\1 1/" \
  -e "s/(PARAMETER ::[[:space:]]*USE_LES[[:space:]]*=)[[:space:]]*[0-9]+/\1 $USE_LES/" \
  -e "s/(PARAMETER ::[[:space:]]*RESIMAX[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $RESIMAX/" \
  -e "s/(PARAMETER ::[[:space:]]*REF_A[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $REF_A/" \
  -e "s/(DOUBLE PRECISION ::[[:space:]]*RE[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $RE/" \
  -e "s/(DOUBLE PRECISION ::[[:space:]]*DTIME[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $DTIME/" \
  "$P"
grep -E "NI|NJ|NK|NDIM|NSWPP|IRESTART|USE_LES|RESIMAX|REF_A|RE =|DTIME =" "$P" | grep -v '^[[:space:]]*!' >&2
cp "$HERE/harness3d.f90" "$SCRATCH/"
cd "$SCRATCH"
FFLAGS="-O2 -ffp-contract=off -I$MPI_INC"
$FLANG $FFLAGS -c comdat_param.f90 comdat_shared.f90
SRCS=$(ls *.f90 | grep -v -e '^comdat_param.f90$' -e '^comdat_shared.f90$' -e '^harness3d.f90$')
$FLANG $FFLAGS -c $SRCS harness3d.f90 2>"$SCRATCH/warn.txt" || { cat "$SCRATCH/warn.txt" >&2; exit 1; }
OUT="$REPO/oracle/_ref/$LABEL"
mkdir -p "$OUT"
$FLANG $FFLAGS -o "$OUT/harness3d" *.o -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
echo "built $OUT/harness3d" >&2
