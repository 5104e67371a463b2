#!/bin/bash
# Build the reference's own Fortran hot path into oracle/_ref/ (checker only).
#
# Compiles, from where they lie under /root/reference, exactly the reference
# sources the CPU path needs, plus oracle/ref_driver.f90 (bind(C) accessors),
# with ROCm flang.  -fdefault-real-8 -fdefault-double-8 is flang's spelling of
# the ifort -r8 used by solver/mix3d.mk:3; -O3 matches mix3d.mk.
# Two libraries are produced because the two compact schemes define the same
# subroutine names (inideriv/derivnper):
#   libemx_ref_c4.so  : solver/derivada.f90            (4th-order Pade, the
#                       scheme of the built binary's CPU path)
#   libemx_ref_c6f.so : src/MSVSingle/derivada_incomp.f90 (6th-order, alpha=1/3)
# Nothing is copied from the reference into the repository.
set -euo pipefail
REF=${EMX_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/solver" ]; then
  echo "build_ref: $REF not present; skipping (prebuilt _ref libraries are used if shipped)" >&2
  exit 0
fi
FC=${FC:-flang}
FFLAGS="-O3 -fPIC -fdefault-real-8 -fdefault-double-8 -w"
mkdir -p "$OUT"
for scheme in c4 c6f; do
  B="$OUT/build_$scheme"
  mkdir -p "$B"
  if [ "$scheme" = c4 ]; then DERIV="$REF/solver/derivada.f90"; else DERIV="$REF/src/MSVSingle/derivada_incomp.f90"; fi
  SRCS=(
    "$REF/src/MSVSingle/flow_mod_incomp2.f90"
    "$REF/solver/multiphase_mod.f90"
    "$DERIV"
    "$REF/solver/deriv_vel.f90"
    "$REF/solver/divergencia.f90"
    "$REF/solver/flujos_compact.f90"
    "$REF/solver/sgdm_incomp.f90"
    "$REF/solver/viscosidad_incomp.f90"
    "$REF/solver/jacob_malla.f90"
    "$REF/src/MSVSingle/herramientas_incomp2.f90"
    "$REF/src/MSVSingle/filtro_incomp.f90"
    "$REF/src/MSVSingle/frontera_y.f90"
    "$REF/src/MSVSingle/frontera_z.f90"
    "$REF/solver/gravar_incomp.f90"
    "$REF/solver/leer_incomp.f90"
    "$HERE/ref_driver.f90"
  )
  OBJS=()
  for s in "${SRCS[@]}"; do
    o="$B/$(basename "${s%.f90}").o"
    (cd "$B" && $FC $FFLAGS -c "$s" -o "$o")
    OBJS+=("$o")
  done
  $FC -shared $FFLAGS "${OBJS[@]}" -o "$OUT/libemx_ref_$scheme.so"
done
echo "build_ref: built $OUT/libemx_ref_c4.so $OUT/libemx_ref_c6f.so"

# Drop-in check program: the reference's own GPU-facing Fortran wrappers
# (deriv_vel_cuda.f90, divergencia_cuda.f90, advectLS.f90), unmodified, with
# emix3d_amd/fortran/emix3d_shim.f90 and oracle/dropin_driver.f90, linked
# against libemix3d_hip.so instead of the reference's CUDA objects.
LIBDIR="$(cd "$HERE/.." && pwd)/emix3d_amd"
if [ -f "$LIBDIR/libemix3d_hip.so" ]; then
  B="$OUT/build_dropin"
  mkdir -p "$B"
  DSRCS=(
    "$REF/src/MSVSingle/flow_mod_incomp2.f90"
    "$REF/solver/multiphase_mod.f90"
    "$REF/solver/deriv_vel_cuda.f90"
    "$REF/solver/divergencia_cuda.f90"
    "$REF/solver/advectLS.f90"
    "$LIBDIR/fortran/emix3d_shim.f90"
    "$HERE/dropin_driver.f90"
  )
  DOBJS=()
  for s in "${DSRCS[@]}"; do
    o="$B/$(basename "${s%.f90}").o"
    (cd "$B" && $FC $FFLAGS -I"$REF/include" -c "$s" -o "$o")
    DOBJS+=("$o")
  done
  $FC $FFLAGS "${DOBJS[@]}" -L"$LIBDIR" -lemix3d_hip \
      -Wl,-rpath,'$ORIGIN/../../emix3d_amd' -o "$OUT/emx_dropin"
  echo "build_ref: built $OUT/emx_dropin"
fi
