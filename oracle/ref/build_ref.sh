#!/bin/sh
# Builds the reference's own Fortran hot path, from its sources where they lie
# under /root/reference, into oracle/_ref/ref_md_<config> (test
# infrastructure: the tests use these programs to pin the C oracle).
#
# Each build preprocesses the reference's sources with a configuration header
# of this directory (control_<config>.h, the reference's compile-time
# switches) and compiles them with amdflang.  Intermediate files (preprocessed
# text, modules, objects) live in a temporary directory that is removed
# afterwards; only the linked programs are written to oracle/_ref/.
#
# -Dk_ring=k_chain: intra_molec.f90:133,272 take the FENE constant of the
# melt chains from k_ring, which init_params.f90:890-900 sets only for
# SYSTEM 4 (rings); for linear chains the macro makes it k_chain.
set -eu
REF=${MFA_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
FC=${FC:-/opt/rocm/bin/amdflang}
[ -d "$REF" ] || { echo "build_ref: $REF absent, nothing built"; exit 0; }
mkdir -p "$OUT"

SOURCES="mfa_common ziggurat R250 INR250 SUN util functions messages \
  make_binning_boxes binning check_skin verlet_positions verlet_velocities \
  dpd_forces lgv_forces ewald_real ewald_k dipolar_correction fluid_fluid \
  fluid_wall intra_molec bending init_force_switch_on init_system init_params \
  gen_brush gen_chain gen_wall gen_droplet conf_default init_config init_obser \
  obser_out observation store_config fix_CM constant_force geom_constraint new_dpd_fd"

# build <config> <output name> <flags>
build() {
  cfg=$1; name=$2; flags=$3
  TMP=$(mktemp -d)
  cp "$HERE/control_$cfg.h" "$TMP/control_simulation.h"
  objs=""
  for f in $SOURCES; do
    cpp -traditional-cpp -P -I"$TMP" -Dk_ring=k_chain < "$REF/$f.f90" > "$TMP/$f.f90"
    (cd "$TMP" && "$FC" -c $flags "$f.f90" -o "$f.o" > "$TMP/$f.log" 2>&1) || {
      cat "$TMP/$f.log"; rm -rf "$TMP"; echo "build_ref: $f failed ($cfg)"; exit 1; }
    objs="$objs $TMP/$f.o"
  done
  cpp -traditional-cpp -P -I"$TMP" < "$HERE/ref_md.F90" > "$TMP/ref_md.f90"
  (cd "$TMP" && "$FC" $flags ref_md.f90 $objs -o "$OUT/$name" > "$TMP/link.log" 2>&1) || {
    cat "$TMP/link.log"; rm -rf "$TMP"; echo "build_ref: link failed ($cfg)"; exit 1; }
  rm -rf "$TMP"
  echo "build_ref: $OUT/$name"
}

# the parity builds: no contraction, so the operation order is the source's
for cfg in brush brush_q brush_bend brush_vv brush_lgv brush_obs melt melt_fso; do
  build "$cfg" "ref_md_$cfg" "-O1 -ffp-contract=off"
done
# a timing build of the brush path (no parity use): -O3 for the host CPU, as
# the reference's users would build it for speed (oracle/calibrate_cpu.py)
build brush ref_md_brush_o3 "-O3 -march=native"
