#!/usr/bin/env bash
# Builds the oracles of the DEM hot path.  TEST INFRASTRUCTURE ONLY.
#
#  oracle/liboracle.so            C restatement (dem3d_oracle.c), gcc, no FMA
#                                 contraction.  Always built.
#  oracle/_ref/libdem3d_ref.so    the reference Fortran modules, compiled from
#                                 their own files under $REF/src with amdflang
#                                 -O2 -ffp-contract=off, plus ref_harness.f90.
#  oracle/_ref/libdem3d_ref_big.so  same, with the two capacity constants of
#                                 MForces raised (MaxNeighbours 50 -> 65536,
#                                 MaxNCell 30000 -> 2147483647) so that systems
#                                 beyond ~20 k particles do not overflow the
#                                 unchecked Neighbour(50) buffer.  The sources
#                                 are streamed through sed into the compiler;
#                                 no copy of them is written anywhere.
#  oracle/_ref/libdem3d_ref_fast.so  capacity-raised as _big, built for speed
#                                 (-O3, contraction on, AVX2/FMA): the CPU
#                                 baseline of the benchmark.  Not bit-exact.
#  oracle/_ref/libdem3d_ref_native.so  the same with -march=native, only when
#                                 DEM3D_REF_NATIVE=1 (timings on this host).
#  oracle/_ref/dem3d_ref          the reference's own program DEM (src/DEM.f90
#                                 over all of its modules), -O2
#                                 -ffp-contract=off.
#  oracle/_ref/dem3d_gpu          the same program, built from the same files
#                                 except that dem3d_amd/fortran/mforces_gpu.f90
#                                 takes the place of src/MForces.f90: its
#                                 ComputeAccelerations runs on the GPU through
#                                 dem3d_amd/libdem3d_hip.so (built first).
#
# The reference part is skipped when $REF/src is absent (e.g. on a machine that
# only runs the prebuilt libraries).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${DEM3D_REFERENCE:-/root/reference}"
CC="${CC:-gcc}"

"$CC" -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -std=c99 \
  -Wall -Wextra -Wno-unused-parameter \
  -o "$HERE/liboracle.so" "$HERE/dem3d_oracle.c" -lm

if [ ! -d "$REF/src" ]; then
  echo "build_ref.sh: $REF/src not found; reference checker not rebuilt" >&2
  exit 0
fi
FC="${FC:-amdflang}"
FFLAGS="-O2 -ffp-contract=off -fPIC -cpp"
MODS="MTypes MTime BinFiles MDimension3D MUtil MCells MConfig MContacts mFluid MForces MDEM"

build_variant() {
  local name="$1" patch="$2" flags="${3:-$FFLAGS}"
  local obj="$HERE/_ref/obj_$name"
  rm -rf "$obj"
  mkdir -p "$obj"
  local objs=()
  for m in $MODS; do
    if [ "$m" = "MForces" ] && [ -n "$patch" ]; then
      sed -e 's/MaxNeighbours = 50/MaxNeighbours = 65536/' \
          -e 's/MaxNCell = 30000_ib/MaxNCell = 2147483647_ib/' "$REF/src/$m.f90" |
        (cd "$obj" && "$FC" $flags -x f95-cpp-input -c - -o "$obj/$m.o")
      grep -q 'MaxNeighbours = 50' "$REF/src/$m.f90"
      grep -q 'MaxNCell = 30000_ib' "$REF/src/$m.f90"
    else
      (cd "$obj" && "$FC" $flags -c "$REF/src/$m.f90" -o "$obj/$m.o")
    fi
    objs+=("$obj/$m.o")
  done
  (cd "$obj" && "$FC" $flags -c "$HERE/ref_harness.f90" -o "$obj/ref_harness.o")
  "$FC" -shared -fPIC -o "$HERE/_ref/libdem3d_ref$name.so" "${objs[@]}" "$obj/ref_harness.o"
  rm -rf "$obj"
}

# The program DEM, as the reference builds it and with the GPU MForces.
build_programs() {
  local obj="$HERE/_ref/obj_prog" flags="-O2 -ffp-contract=off -cpp"
  local root lib
  root="$(cd "$HERE/.." && pwd)"
  lib="$root/dem3d_amd/libdem3d_hip.so"
  rm -rf "$obj"
  mkdir -p "$obj/common" "$obj/ref" "$obj/gpu"
  local common=()
  for m in MTypes MTime BinFiles MDimension3D MUtil MCells MConfig MContacts mFluid; do
    "$FC" $flags -module-dir "$obj/common" -c "$REF/src/$m.f90" -o "$obj/common/$m.o"
    common+=("$obj/common/$m.o")
  done
  for m in MForces MDEM DEM; do
    "$FC" $flags -module-dir "$obj/ref" -I "$obj/common" -c "$REF/src/$m.f90" -o "$obj/ref/$m.o"
  done
  "$FC" -o "$HERE/_ref/dem3d_ref" "${common[@]}" "$obj/ref/MForces.o" "$obj/ref/MDEM.o" \
    "$obj/ref/DEM.o"
  if [ -f "$lib" ]; then
    "$FC" $flags -module-dir "$obj/gpu" -c "$root/dem3d_amd/fortran/dem3d_gpu.f90" \
      -o "$obj/gpu/dem3d_gpu.o"
    "$FC" $flags -module-dir "$obj/gpu" -I "$obj/common" -c \
      "$root/dem3d_amd/fortran/mforces_gpu.f90" -o "$obj/gpu/MForces.o"
    for m in MDEM DEM; do
      "$FC" $flags -module-dir "$obj/gpu" -I "$obj/common" -c "$REF/src/$m.f90" \
        -o "$obj/gpu/$m.o"
    done
    "$FC" -o "$HERE/_ref/dem3d_gpu" "${common[@]}" "$obj/gpu/dem3d_gpu.o" "$obj/gpu/MForces.o" \
      "$obj/gpu/MDEM.o" "$obj/gpu/DEM.o" "$lib" -Wl,-rpath,'$ORIGIN/../../dem3d_amd'
  fi
  rm -rf "$obj"
}

mkdir -p "$HERE/_ref"
build_programs
build_variant "" ""
build_variant "_big" "yes"
# the reference's own optimisation (Makefile: -O3 -march=native, contraction
# left on) for timing only; -march=x86-64-v3 (AVX2 + FMA) instead of native so
# that the library also runs on the GPU machine's host CPU
build_variant "_fast" "yes" "-O3 -march=x86-64-v3 -fPIC -cpp"
if [ "${DEM3D_REF_NATIVE:-0}" = "1" ]; then
  build_variant "_native" "yes" "-O3 -march=native -fPIC -cpp"
fi
echo "build_ref.sh: built $HERE/_ref/libdem3d_ref.so, libdem3d_ref_big.so, libdem3d_ref_fast.so, dem3d_ref and dem3d_gpu"
