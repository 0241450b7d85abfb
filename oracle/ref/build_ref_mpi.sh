#!/usr/bin/env bash
# Build the reference 3D LES solver (LES/v2.3_LES-Reynolds stresses) as the reference's own
# program -- its main.f90, MPI domain decomposition and console output unchanged -- for the CPU
# baseline: mpirun -np NDIMJ*NDIMK on the host cores, the (J, K) block decomposition of
# define.f90:18-38.
#
# The sources are copied to a scratch directory outside the repository; only compile-time
# constants of comdat_param.f90 are substituted: grid size, NDIMJ x NDIMK ranks, a cold start
# (IRESTART = 0: uniform flow, no restart file), Re, DTIME, and RTIME_MAX = RTIME_END = NSTEP *
# DTIME so the time loop ends after NSTEP steps.  NSWPP and RESIMAX stay the reference's own.
# The binary (never committed) is written to oracle/_ref/<label>/rles.
#
# usage: build_ref_mpi.sh LABEL NI NJ NK NDIMJ NDIMK RE DTIME NSTEP
set -euo pipefail
REF_SRC="/root/reference/LES/v2.3_LES-Reynolds stresses"
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
LABEL=$1; NI=$2; NJ=$3; NK=$4; NDJ=$5; NDK=$6; RE=$7; DTIME=$8; NSTEP=$9
FLANG=${FLANG:-/opt/rocm/lib/llvm/bin/flang}
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}

if [ ! -d "$REF_SRC" ]; then
  echo "reference sources not present; oracle/_ref cannot be built here" >&2
  exit 3
fi
# the end time, a hair below NSTEP * DTIME (the loop ends once RTIME >= RTIME_MAX)
TEND=$(python3 -c "print(('%.16e' % ($NSTEP * float('$DTIME'.replace('D', 'e')) * (1 - 1e-9))).replace('e', 'D'))")
SCRATCH=$(mktemp -d /tmp/fvm_refmpi.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
cp "$REF_SRC"/*.f90 "$SCRATCH"/
P="$SCRATCH/comdat_param.f90"
sed -E -i \
  -e "s/(PARAMETER ::[[:space:]]*NI[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NI/" \
  -e "s/(PARAMETER ::[[:space:]]*NJ[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NJ/" \
  -e "s/(PARAMETER ::[[:space:]]*NK[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NK/" \
  -e "s/(PARAMETER ::[[:space:]]*NDIMJ[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NDJ/" \
  -e "s/(PARAMETER ::[[:space:]]*NDIMK[[:space:]]*=)[[:space:]]*[0-9]+/\1 $NDK/" \
  -e "s/(PARAMETER ::[[:space:]]*IRESTART[[:space:]]*=)[[:space:]]*[0-9]+/\1 0/" \
  -e "s/(DOUBLE PRECISION ::[[:space:]]*RE[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $RE/" \
  -e "s/(DOUBLE PRECISION ::[[:space:]]*DTIME[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $DTIME/" \
  -e "s/(DOUBLE PRECISION ::[[:space:]]*RTIME_MAX[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $TEND/" \
  -e "s/(DOUBLE PRECISION, PARAMETER ::[[:space:]]*RTIME_END[[:space:]]*=)[[:space:]]*[-+.0-9D]+/\1 $TEND/" \
  "$P"
grep -E "NI|NJ|NK|NDIM|NSWPP|IRESTART|RESIMAX|RE =|DTIME =|RTIME_" "$P" | grep -v '^[[:space:]]*!' >&2
cd "$SCRATCH"
FFLAGS="-O2 -I$MPI_INC"
$FLANG $FFLAGS -c comdat_param.f90 comdat_shared.f90
SRCS=$(ls *.f90 | grep -v -e '^comdat_param.f90$' -e '^comdat_shared.f90$')
$FLANG $FFLAGS -c $SRCS 2>"$SCRATCH/warn.txt" || { cat "$SCRATCH/warn.txt" >&2; exit 1; }
OUT="$REPO/oracle/_ref/$LABEL"
mkdir -p "$OUT"
$FLANG $FFLAGS -o "$OUT/rles" *.o -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
echo "built $OUT/rles" >&2
