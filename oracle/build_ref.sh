#!/usr/bin/env bash
# Builds the spray-simulation reference with flang, linked to the project's
# golden-vector driver (oracle/ref/fvk_ref_driver.f90).  TEST INFRASTRUCTURE,
# run in the development container only (the reference is not on GPU hosts).
#
# Objects and module files go to a scratch directory outside the repository;
# only the driver executables land in oracle/_ref/:
#   _ref/shipped/fvk_ref_driver   reference sources compiled where they lie
#   _ref/relaxed/fvk_ref_driver   same, except that cfd_setup.f90 is taken
#       from a scratch copy with two hard-coded locals switched:
#         simpler=.true.  -> simpler=.false.   (cfd_setup.f90:581; enables the
#                                                relaxed update at :675-683)
#         call solv_cg(1, -> call solv_cg(0,   (cfd_setup.f90:662,742; Jacobi)
#       which is the SIMPLE/Jacobi-PCG configuration the reference cannot
#       select without editing a literal;
#   _ref/cg/fvk_ref_cg            the reference's solv_cg with a timing
#       driver (oracle/ref/fvk_ref_cg.f90), the CPU baseline of bench.py;
#   _ref/outp/fvk_ref_outp        the reference's Tecplot writers
#       (cfd_outp.f90) behind oracle/ref/fvk_ref_outp.f90;
#   _ref/gmsh/gmsh_ref            flowlab's gmsh reader and field writers
#       (flowlab/src/mesh.cpp) behind oracle/ref/gmsh_ref.cpp;
#   _ref/flowlab/flowlab_ref_driver  flowlab's serial CSR SIMPLE (its
#       Fortran modules) behind oracle/ref/flowlab_ref_driver.f90.
set -euo pipefail
REF=${FVK_REFERENCE_DIR:-/root/reference/spray-simulation}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-flang}
FFLAGS=${FFLAGS:--O2}
MODS="precision solv_data solv_res solv_precond solv_direct solv_bicg solv_cg
      mom_data mom_math mom_closure mom_solve cfd_data grd_data grd_solv grd_math
      grd_constr grd_read grd_conn grd_geom cfd_constr spr_data spr_inject
      spr_impact spr_hydro solv_conv cfd_define cfd_setup"
[ -d "$REF" ] || { echo "reference not found at $REF" >&2; exit 1; }
SCRATCH=$(mktemp -d /tmp/fvk_ref_build.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT

build() {  # $1 variant name, $2 directory holding cfd_setup.f90
  local v=$1 setup_dir=$2 w=$SCRATCH/$1 objs=""
  mkdir -p "$w" "$OUT/$v"
  for m in $MODS; do
    local src=$REF/$m.f90
    [ "$m" = cfd_setup ] && src=$setup_dir/cfd_setup.f90
    (cd "$w" && $FC $FFLAGS -c "$src" -o "$m.o")
    objs="$objs $w/$m.o"
  done
  (cd "$w" && $FC $FFLAGS -c "$HERE/ref/fvk_ref_driver.f90" -o driver.o)
  $FC $FFLAGS -o "$OUT/$v/fvk_ref_driver" "$w/driver.o" $objs
}

build shipped "$REF"
mkdir -p "$SCRATCH/relaxed_src"
sed -e 's/simpler=\.true\./simpler=.false./' -e 's/call solv_cg(1,/call solv_cg(0,/' \
    "$REF/cfd_setup.f90" > "$SCRATCH/relaxed_src/cfd_setup.f90"
[ "$(grep -c 'simpler=.false.' "$SCRATCH/relaxed_src/cfd_setup.f90")" -ge 2 ]
[ "$(grep -c 'call solv_cg(0,' "$SCRATCH/relaxed_src/cfd_setup.f90")" -eq 2 ]
build relaxed "$SCRATCH/relaxed_src"

# the timing driver of the reference's PCG (bench.py's CPU baseline): the
# solver modules only, compiled where they lie
w=$SCRATCH/cg objs=""
mkdir -p "$w" "$OUT/cg"
for m in precision solv_data solv_res solv_precond solv_direct solv_cg; do
  (cd "$w" && $FC $FFLAGS -c "$REF/$m.f90" -o "$m.o")
  objs="$objs $w/$m.o"
done
(cd "$w" && $FC $FFLAGS -c "$HERE/ref/fvk_ref_cg.f90" -o driver.o)
$FC $FFLAGS -o "$OUT/cg/fvk_ref_cg" "$w/driver.o" $objs
# the reference's Tecplot writers (cfd_outp.f90) behind a driver of this
# project (oracle/ref/fvk_ref_outp.f90): golden output files, shipped
# configuration
w=$SCRATCH/outp objs=""
mkdir -p "$w" "$OUT/outp"
for m in $MODS cfd_outp; do
  (cd "$w" && $FC $FFLAGS -c "$REF/$m.f90" -o "$m.o")
  objs="$objs $w/$m.o"
done
(cd "$w" && $FC $FFLAGS -c "$HERE/ref/fvk_ref_outp.f90" -o driver.o)
$FC $FFLAGS -o "$OUT/outp/fvk_ref_outp" "$w/driver.o" $objs
# flowlab's gmsh reader and field writers (flowlab/src/mesh.cpp, compiled
# where it lies) behind a driver of this project (oracle/ref/gmsh_ref.cpp)
FLOWLAB=${FVK_FLOWLAB_DIR:-$REF/../flowlab/src}
mkdir -p "$OUT/gmsh"
g++ -O2 -I"$FLOWLAB" -o "$OUT/gmsh/gmsh_ref" "$HERE/ref/gmsh_ref.cpp" "$FLOWLAB/mesh.cpp"
# flowlab's serial CSR SIMPLE: its Fortran modules compiled where they lie
# (-cpp, no USE_MPI: the serial build) behind a driver of this project
# (oracle/ref/flowlab_ref_driver.f90)
w=$SCRATCH/flowlab objs=""
mkdir -p "$w" "$OUT/flowlab"
for m in prec_m constants_m vector_utils_m matrix_utils_m graph_m hash_m solv_m geometry_m transport_m; do
  (cd "$w" && $FC $FFLAGS -cpp -c "$FLOWLAB/$m.f90" -o "$m.o")
  objs="$objs $w/$m.o"
done
(cd "$w" && $FC $FFLAGS -c "$HERE/ref/flowlab_ref_driver.f90" -o driver.o)
$FC $FFLAGS -o "$OUT/flowlab/flowlab_ref_driver" "$w/driver.o" $objs
echo "built $OUT/shipped/fvk_ref_driver $OUT/relaxed/fvk_ref_driver $OUT/cg/fvk_ref_cg $OUT/outp/fvk_ref_outp" \
     "$OUT/gmsh/gmsh_ref $OUT/flowlab/flowlab_ref_driver"
