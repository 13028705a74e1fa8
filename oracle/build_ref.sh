#!/bin/bash
# Builds the reference's inlet Zou-He (Boundary::ApplyBc), D3Q19 lattice
# tables (LatticeModel) and STL readers (Reader) from its own sources into
# oracle/_ref/ (test infrastructure; never shipped).  Needs the reference
# checkout (default /root/reference) and the MPI headers in the image (include.h includes mpi.h;
# nothing in these files calls MPI, so no MPI library is linked).
# Compiled files: Utility.cpp LatticeModel.cpp VelocitySet.cpp Node.cpp Boundary.cpp GeometryReader.cpp.
# Not built: Simulation.cpp / OutputFile.cpp (need the VTK library, absent) and
# ComplexGeometry.cpp (includes mpi_call.h, rejected by gcc and clang at :143).
set -e
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MPI_INC=${MPI_INC:-/opt/conda/include}
[ -f "$REF/Boundary.cpp" ] || { echo "reference not found at $REF"; exit 3; }
[ -f "$MPI_INC/mpi.h" ] || { echo "mpi.h not found"; exit 3; }
mkdir -p "$OUT"
g++ -std=c++17 -O2 -ffp-contract=off -fopenmp -fpermissive -w -I"$REF" -I"$MPI_INC" \
    "$HERE/ref_driver.cpp" \
 "$REF/Utility.cpp" "$REF/LatticeModel.cpp" "$REF/VelocitySet.cpp" "$REF/Node.cpp" \
    "$REF/Boundary.cpp" "$REF/GeometryReader.cpp" \
    -o "$OUT/ref_driver"
echo "built $OUT/ref_driver"
