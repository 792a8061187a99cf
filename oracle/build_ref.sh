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
# Three builds: the default FP64 one (oracle/_ref/libmpdata_ref.so), the
# reference's own single-precision switch -DFLOAT_PRECISION=4
# (common/precisions.f90:9-17: euwp = REAL32) as
# oracle/_ref/libmpdata_ref32.so, and oracle/_ref/libmpdata_ref_dropin.so:
# the FP64 build with the reference's three stage sources (upwind3d-gpubc,
# antidiff3d-gauge-gpubc, antidiff3d-standard-gpubc) replaced by this
# repository's literal drop-in modules of the same names
# (advectionhpctester_amd/fortran/mpdata_dropin.F90, over libmpdata_hip.so),
# so that the reference's own advec.f90 compiles its procedure-pointer
# assignments against them and its advec_dwarf runs the HIP kernels.
# Intermediate files (preprocessed sources, objects, modules) live in
# temporary directories that are removed afterwards; the libraries are the
# only outputs.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-amdflang}
FFLAGS="-O2 -fPIC"

if [ ! -d "$REF/common" ]; then
    echo "build_ref: reference tree not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
WORKS=()
trap 'rm -rf "${WORKS[@]}"' EXIT

# build_one <extra defines> <library name> [dropin]
build_one() {
    local DEFS="-DNOCUDACODE -DTIMERSCPU $1" LIB="$2" DROPIN="${3:-}"
    local WORK
    WORK=$(mktemp -d "${TMPDIR:-/tmp}/mpdata_ref.XXXXXX")
    WORKS+=("$WORK")
    mkdir -p "$WORK/pp" "$WORK/obj"
    (
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
    if [ -n "$DROPIN" ]; then
        local FD="$HERE/../advectionhpctester_amd/fortran"
        $FC $FFLAGS -cpp -DFLOAT_PRECISION=8 -c "$FD/mpdata_hip.F90" -o mpdata_hip.o
        $FC $FFLAGS -cpp -DFLOAT_PRECISION=8 -c "$FD/mpdata_dropin.F90" -o mpdata_dropin.o
    else
        pp advection/src_algorithms/upwind3d-gpubc.f90
        pp advection/src_algorithms/antidiff3d-gauge-gpubc.f90
        pp advection/src_algorithms/antidiff3d-standard-gpubc.f90
    fi
    pp advection/src_algorithms/module_velprd.f90
    pp advection/src_algorithms/velprd_driver.f90
    pp advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90
    pp advection/src_drivers/advec.f90
    pp advection/src_testdata/advection_test_setup.f90
    $FC $FFLAGS -c "$HERE/ref_shim.f90" -o ref_shim.o
    if [ -n "$DROPIN" ]; then
        $FC -shared -fPIC -o "$OUT/$LIB" ./*.o -L"$HERE/../advectionhpctester_amd" -lmpdata_hip \
            -Wl,-rpath,'$ORIGIN/../../advectionhpctester_amd'
    else
        $FC -shared -fPIC -o "$OUT/$LIB" ./*.o
    fi
    )
    echo "build_ref: wrote $OUT/$LIB"
}

build_one "" libmpdata_ref.so
build_one "-DFLOAT_PRECISION=4" libmpdata_ref32.so
if [ -f "$HERE/../advectionhpctester_amd/libmpdata_hip.so" ]; then
    build_one "" libmpdata_ref_dropin.so dropin
else
    echo "build_ref: advectionhpctester_amd/libmpdata_hip.so not built; no drop-in build" >&2
fi
