#!/usr/bin/env bash
# Build the reference Fortran hot path into oracle/_ref/ (test infrastructure).
#
# Compiles the reference's own sources where they lie under /root/reference
# (nothing is copied into this repository) together with oracle/ref_harness.f90,
# with amdflang and the MPICH headers/libraries shipped in /opt/conda.
# -ffp-contract=off keeps every floating-point operation separately rounded,
# as the C oracle (oracle/rt6d_oracle.c) is compiled, so the two can be
# compared bit for bit.  diagnose.f90 and the netCDF writers are not built:
# they need netcdf.inc, which the image does not have.
#
# Output: oracle/_ref/ref_harness (binary), oracle/_ref/*.o *.mod.
# Exits 0 without building when /root/reference is absent (the GPU box).
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF=${RT6D_REFERENCE:-/root/reference}/for_src
OUT="$here/_ref"
if [ ! -d "$REF" ]; then
  echo "build_ref: $REF not present; skipping reference build"
  exit 0
fi
FC=${FC:-amdflang}
MPIINC=${MPIINC:-/opt/conda/include}
MPILIB=${MPILIB:-/opt/conda/lib}
FFLAGS="-O2 -ffp-contract=off -w"
mkdir -p "$OUT"
cd "$OUT"
for f in timing_module main_module parallel_mpi calc_grid calc_parameter superbee_fluxes integrate; do
  "$FC" $FFLAGS -I"$MPIINC" -c "$REF/$f.f90" -o "$f.o"
done
"$FC" $FFLAGS -I"$MPIINC" -c "$here/ref_harness.f90" -o ref_harness.o
"$FC" $FFLAGS -o ref_harness ref_harness.o timing_module.o main_module.o parallel_mpi.o \
  calc_grid.o calc_parameter.o superbee_fluxes.o integrate.o \
  -L"$MPILIB" -Wl,-rpath,"$MPILIB" -lmpifort -lmpi
echo "build_ref: built $OUT/ref_harness"
