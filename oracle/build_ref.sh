#!/usr/bin/env bash
# Builds oracle/_ref/libmpdata_ref.so from the reference's own Fortran
# sources (CPU path) plus oracle/ref_shim.f90.  TEST INFRASTRUCTURE ONLY.
#
# Recipe: the reference's CMake preprocesses most sources with
# "cpp -E -P | sed" so that macro-generated one-line loops are split into
# statements (CMakeLists.txt:157-205); the other sources are compiled
# directly with -cpp.  Flags follow a CPU, non-MPI configuration
# (-DNOCUDACODE -DTIMERSCPU, CMakeLists.txt:14-47) at -O2 without -Ofast so
# that signed zeros, division and evaluation order follow IEEE rules.
# Intermediate files (preprocessed sources, objects, modules) live in a
# temporary directory that is removed afterwards; the only output is
# oracle/_ref/libmpdata_ref.so.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-amdflang}
FFLAGS="-O2 -fPIC"
DEFS="-DNOCUDACODE -DTIMERSCPU"

if [ ! -d "$REF/common" ]; then
    echo "build_ref: reference tree not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
WORK=$(mktemp -d "${TMPDIR:-/tmp}/mpdata_ref.XXXXXX")
trap 'rm -rf "$WORK"' EXIT
mkdir -p "$WORK/pp" "$WORK/obj"
cd "$WORK/obj"

pp() {   # pp <source> : preprocess + split like the reference's CMake recipe
    local src="$REF/$1" dst="$WORK/pp/$(basename "$1")"
    cpp -traditional-cpp -P $DEFS -I"$(dirname "$src")" "$src" |
        sed -e 's/&/\&\n/g' -e 's/;/;\n/g' -e 's/>>>/>>>\n/g' -e '/^$/d' > "$dst"
    $FC $FFLAGS -c "$dst" -o "$(basename "$1").o"
}
direct() {   # direct <source> : compiled as-is with the C preprocessor
    $FC $FFLAGS -cpp $DEFS -c "$REF/$1" -o "$(basename "$1").o"
}

direct common/precisions.f90
direct common/epsilons.f90
direct common/eulag_lib_parameters.f90
pp common/scratch_datafields_modular.f90
pp common/mpi_parallel.f90
direct common/noise.f90
direct common/bconditions.f90
pp common/geometry_datafields_modular.f90
pp common/geometry.f90
pp common/diagutils.f90
pp common/diagnostics.f90
pp advection/src_algorithms/upwind3d-gpubc.f90
pp advection/src_algorithms/antidiff3d-gauge-gpubc.f90
pp advection/src_algorithms/antidiff3d-standard-gpubc.f90
pp advection/src_algorithms/module_velprd.f90
pp advection/src_algorithms/velprd_driver.f90
pp advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90
pp advection/src_drivers/advec.f90
pp advection/src_testdata/advection_test_setup.f90
$FC $FFLAGS -c "$HERE/ref_shim.f90" -o ref_shim.o
$FC -shared -fPIC -o "$OUT/libmpdata_ref.so" ./*.o
echo "build_ref: wrote $OUT/libmpdata_ref.so"
