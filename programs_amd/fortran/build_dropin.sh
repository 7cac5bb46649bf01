#!/usr/bin/env bash
# Builds the Fortran drop-in demonstration: the reference's own host code
# (grd_read / grd_conn / grd_geom, data and constructor modules) and the
# golden-vector driver oracle/ref/fvk_ref_driver.f90, with the hot-path
# modules (solv_res_m, solv_precond_m, solv_direct_m, solv_bicg_m, solv_cg_m,
# grd_solv_m, cfd_setup_m) replaced by the ISO_C_BINDING modules of this
# directory that call libfvk.so.  Development container only (needs
# /root/reference and flang); the executable lands in
# oracle/_ref/dropin/fvk_ref_driver because it contains reference code.
set -euo pipefail
REF=${FVK_REFERENCE_DIR:-/root/reference/spray-simulation}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT=$ROOT/oracle/_ref/dropin
FC=${FC:-flang}
FFLAGS=${FFLAGS:--O2}
[ -d "$REF" ] || { echo "reference not found at $REF" >&2; exit 1; }
[ -f "$ROOT/programs_amd/libfvk.so" ] || { echo "build libfvk.so first" >&2; exit 1; }
SCRATCH=$(mktemp -d /tmp/fvk_dropin_build.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
mkdir -p "$OUT"
objs=""
unit() {  # compile one source into the scratch directory
  local src=$1 name
  name=$(basename "$src" .f90)
  (cd "$SCRATCH" && $FC $FFLAGS -c "$src" -o "$name.o")
  objs="$objs $SCRATCH/$name.o"
}
for m in precision solv_data; do unit "$REF/$m.f90"; done
unit "$HERE/fvk_c.f90"
unit "$HERE/fvk_solv.f90"
for m in mom_data mom_math mom_closure mom_solve cfd_data grd_data; do unit "$REF/$m.f90"; done
unit "$HERE/fvk_grd_solv.f90"
for m in grd_math grd_constr grd_read grd_conn grd_geom cfd_constr spr_data spr_inject spr_impact \
         spr_hydro solv_conv cfd_define; do unit "$REF/$m.f90"; done
unit "$HERE/fvk_cfd_setup.f90"
unit "$ROOT/oracle/ref/fvk_ref_driver.f90"
$FC $FFLAGS -o "$OUT/fvk_ref_driver" $objs -L"$ROOT/programs_amd" -lfvk \
    -Wl,-rpath,'$ORIGIN/../../../programs_amd'
# stamp: the hash of the sources this executable was built from; the GPU test
# refuses an executable whose stamp differs from the tree it runs in
(cd "$ROOT" && sha256sum programs_amd/fortran/fvk_c.f90 programs_amd/fortran/fvk_solv.f90 \
   programs_amd/fortran/fvk_grd_solv.f90 programs_amd/fortran/fvk_cfd_setup.f90 \
   oracle/ref/fvk_ref_driver.f90 | sha256sum | cut -d' ' -f1) > "$OUT/sources.sha256"
echo "built $OUT/fvk_ref_driver"

# flowlab: its own Fortran modules with solv_m replaced by fvk_fl_solv.f90
# (lin_eqns_solver on the GPU), behind the golden-vector driver
# oracle/ref/flowlab_ref_driver.f90 -> oracle/_ref/dropin_flowlab/
FLOWLAB=${FVK_FLOWLAB_DIR:-$REF/../flowlab/src}
OUTF=$ROOT/oracle/_ref/dropin_flowlab
mkdir -p "$OUTF" "$SCRATCH/flowlab"
fobjs=""
funit() {
  local src=$1 name
  name=$(basename "$src" .f90)
  (cd "$SCRATCH/flowlab" && $FC $FFLAGS -cpp -c "$src" -o "$name.o")
  fobjs="$fobjs $SCRATCH/flowlab/$name.o"
}
for m in prec_m constants_m vector_utils_m matrix_utils_m graph_m hash_m; do funit "$FLOWLAB/$m.f90"; done
funit "$HERE/fvk_fl_solv.f90"
for m in geometry_m transport_m; do funit "$FLOWLAB/$m.f90"; done
funit "$ROOT/oracle/ref/flowlab_ref_driver.f90"
$FC $FFLAGS -o "$OUTF/flowlab_ref_driver" $fobjs -L"$ROOT/programs_amd" -lfvk \
    -Wl,-rpath,'$ORIGIN/../../../programs_amd'
(cd "$ROOT" && sha256sum programs_amd/fortran/fvk_fl_solv.f90 oracle/ref/flowlab_ref_driver.f90 \
   | sha256sum | cut -d' ' -f1) > "$OUTF/sources.sha256"
echo "built $OUTF/flowlab_ref_driver"
