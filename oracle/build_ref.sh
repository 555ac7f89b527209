#!/bin/bash
# Test infrastructure only.
#
# Builds the reference NEXD-3D *solver* from its own Fortran sources where
# they lie (/root/reference/src), for one polynomial order and one real kind,
# into oracle/_ref/:
#   oracle/_ref/o<N>_<f64|f32>/solver      the reference's program_solver
#   oracle/_ref/o<N>_<f64|f32>/movie       the reference's program_movie
#   oracle/_ref/o<N>_<f64|f32>/ref_probe   oracle/ref_probe.f90 (operator dump)
#   oracle/_ref/o<N>_<f64|f32>/ref_geom_probe  oracle/ref_geom_probe.f90
#                                          (geometry, location, attenuation)
#   oracle/_ref/o<N>_<f64|f32>/libnexdref.so  the reference modules
#
# Compile-time configuration is the reference's own mechanism: the two
# parameters CUSTOM_REAL and ORDER of src/constants.h.  The module that
# includes it is regenerated with those two values in a scratch directory
# outside the repository; nothing of the reference is written into the repo.
#
# Toolchain: amdflang (ROCm), MKL (LAPACK ?getrf/?getri) and MPICH from
# /opt/conda, all present in the image.  METIS (needed only by the mesher's
# createMesh, src/meshMod.f90:422) is absent and is NOT replaced: the modules
# go into a shared library whose one unresolved symbol is never called by the
# solver path, and the solver executable is linked with
# --allow-shlib-undefined.  The mesher is not built.
#
# With a third argument "gpu" the build is the reference solver whose time
# loop calls libnexd_gpu.so: the fragments of
# nexd-3d_amd/fortran/timeloop3d_gpu_patch.f90 are spliced into the scratch
# copy of src/timeloopMod.f90 (oracle/splice_timeloop.py), the modules are
# compiled with nexd_gpu_iface, and the solver links the library, into
# oracle/_ref/o<N>_f64_gpu/.
#
# usage: oracle/build_ref.sh <order> <f64|f32> [gpu]
set -euo pipefail
ORDER=${1:?order}
PREC=${2:-f64}
VARIANT=${3:-}
REF=${NEXD_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref/o${ORDER}_${PREC}${VARIANT:+_$VARIANT}
PKG=$HERE/../nexd-3d_amd
SCRATCH=$(mktemp -d /tmp/nexdref_o${ORDER}_${PREC}.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
mkdir -p "$OUT"
[ -d "$REF/src" ] || { echo "reference sources not found at $REF/src" >&2; exit 2; }

case $PREC in
  f64) KIND=SIZE_DOUBLE ;;
  f32) KIND=SIZE_REAL ;;
  *) echo "precision must be f64 or f32" >&2; exit 2 ;;
esac

# constants.h with the two compile-time parameters set, and the one-line
# module that includes it.
sed -e "s/^\(\s*integer, parameter :: CUSTOM_REAL = \).*/\1${KIND}/" \
    -e "s/^\(\s*integer, parameter :: ORDER = \).*/\1${ORDER}/" \
    "$REF/src/constants.h" > "$SCRATCH/constants.h"
grep -q "CUSTOM_REAL = ${KIND}" "$SCRATCH/constants.h"
grep -q "ORDER = ${ORDER}\$" "$SCRATCH/constants.h"
printf 'module constantsMod\n  include "constants.h"\nend module constantsMod\n' > "$SCRATCH/constantsMod.f90"

# readMeshVar reads qp and qs (src/meshMod.f90:1608-1609) into arrays it never
# allocates (allocation list src/meshMod.f90:1554-1575).  gfortran lets that
# pass; flang's runtime aborts with "Bad address for I/O item".  The scratch
# copy of meshMod.f90 gets the two missing allocations inside readMeshVar
# only; the solver never uses qp/qs.  This is the one source deviation.
awk '
  /subroutine readMeshVar/ { inread = 1 }
  /end subroutine readMeshVar/ { inread = 0 }
  { print }
  inread && /allocate\(this%vpu\(this%nelem\),this%vsu\(this%nelem\),this%muu\(this%nelem\),this%lambdau\(this%nelem\)\)/ {
    print "        allocate(this%qp(this%nelem),this%qs(this%nelem))"
  }' "$REF/src/meshMod.f90" > "$SCRATCH/meshMod.f90"
[ "$(grep -c 'allocate(this%qp(this%nelem),this%qs(this%nelem))' "$SCRATCH/meshMod.f90")" = 1 ]

FC=${FC:-amdflang}
FFLAGS="-O2 -fPIC -ffixed-line-length=none -I$SCRATCH -I/opt/conda/include -module-dir $SCRATCH"
GPUOBJ=""
GPULIB=""
if [ "$VARIANT" = gpu ]; then
  [ "$PREC" = f64 ] || { echo "the library takes REAL(8): gpu builds are f64" >&2; exit 2; }
  [ -f "$PKG/libnexd_gpu.so" ] || { echo "build libnexd_gpu.so first (make -C nexd-3d_amd)" >&2; exit 2; }
  python3 "$HERE/splice_timeloop.py" "$REF/src/timeloopMod.f90" "$PKG/fortran/timeloop3d_gpu_patch.f90" \
      "$SCRATCH/timeloopMod.f90"
  $FC $FFLAGS -c "$PKG/fortran/nexd_gpu_iface.f90" -o "$SCRATCH/nexd_gpu_iface.o"
  GPUOBJ="$SCRATCH/nexd_gpu_iface.o"
  GPULIB="-L$PKG -lnexd_gpu -Wl,-rpath,\$ORIGIN/../../../nexd-3d_amd"
fi
MODS="mpiincludeMod typeMod fileParameterMod parameterMod logoMod allocateMod mpiMod
      rosettaGammaMod jacobiMod matrixMod linearSystemMod tetTrafoMod simplexMod
      vandermondeMod gllMod warpfactorMod nodesMod dmatricesMod liftMod
      geometricFactorsMod normalsMod stfMod sourceReceiverMod anelasticMod waveMod
      tetraNeighborMod vtkMod plotMod pmlMod rungekuttaMod meshMod timeloopMod
      collectMovieMod"
OBJS="$SCRATCH/constantsMod.o $GPUOBJ"
$FC $FFLAGS -c "$SCRATCH/constantsMod.f90" -o "$SCRATCH/constantsMod.o"
for m in $MODS; do
  srcf="$REF/src/$m.f90"
  [ "$m" = meshMod ] && srcf="$SCRATCH/meshMod.f90"
  [ "$m" = timeloopMod ] && [ "$VARIANT" = gpu ] && srcf="$SCRATCH/timeloopMod.f90"
  $FC $FFLAGS -c "$srcf" -o "$SCRATCH/$m.o" 2> "$SCRATCH/$m.log" || { cat "$SCRATCH/$m.log"; exit 1; }
  OBJS="$OBJS $SCRATCH/$m.o"
done
LIBS="-L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpifort -lmpi -lmkl_rt"
$FC -shared -o "$OUT/libnexdref.so" $OBJS $LIBS $GPULIB
$FC $FFLAGS -c "$REF/src/program_solver.f90" -o "$SCRATCH/program_solver.o"
$FC -o "$OUT/solver" "$SCRATCH/program_solver.o" -L"$OUT" -Wl,-rpath,'$ORIGIN' -lnexdref $LIBS $GPULIB \
    -Wl,--allow-shlib-undefined
$FC $FFLAGS -c "$REF/src/program_movie.f90" -o "$SCRATCH/program_movie.o"
$FC -o "$OUT/movie" "$SCRATCH/program_movie.o" -L"$OUT" -Wl,-rpath,'$ORIGIN' -lnexdref $LIBS $GPULIB \
    -Wl,--allow-shlib-undefined
$FC $FFLAGS -c "$HERE/ref_probe.f90" -o "$SCRATCH/ref_probe.o"
$FC -o "$OUT/ref_probe" "$SCRATCH/ref_probe.o" -L"$OUT" -Wl,-rpath,'$ORIGIN' -lnexdref $LIBS -Wl,--allow-shlib-undefined
$FC $FFLAGS -c "$HERE/ref_geom_probe.f90" -o "$SCRATCH/ref_geom_probe.o"
$FC -o "$OUT/ref_geom_probe" "$SCRATCH/ref_geom_probe.o" -L"$OUT" -Wl,-rpath,'$ORIGIN' -lnexdref $LIBS \
    -Wl,--allow-shlib-undefined
echo "built $OUT"
