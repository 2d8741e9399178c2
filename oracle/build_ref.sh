#!/usr/bin/env bash
# Builds oracle/_ref/ from the reference's own CPU sources, read where they lie
# under /root/reference (nothing is copied into this repository).  Test
# infrastructure only: the outputs are checkers and the reference CPU baseline.
#
#   libpvref.so       -O3 -ffp-contract=off: bit-exact parity checker
#   libpvref_fast.so  the reference's own host flags (PV/Makefile:45:
#                     -O3 -march=native -funroll-loops --fast-math), with
#                     -march=x86-64-v3 so the library runs on any AVX2 host
#   libpvref_o3.so    the same without --fast-math (-O3 -march=x86-64-v3
#                     -funroll-loops): the second CPU baseline
#
# Units built: PV/AmdLib/*.c, PV/006_get_dependencies_GLU.cpp,
# PV/007_convert_dependencies_in_levels.cpp,
# PV/008_convert_chains_columns_levels_into_vector.cpp,
# PV/009_get_max_numbers_per_column.cpp, PV/GLU_CSC_CPU_levels_strings.cpp,
# PV/GLU_CSC_CPU_Triangular.cpp, plus oracle/ref_harness.cpp.
# Not built: every .cu (no nvcc) and every unit that needs PV/util.cpp
# (it includes cuda.h): 004a/004b/004c/010/011/002/002b/003.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="${PVGLU_REFERENCE:-/root/reference}/PVPP_Plant_Simulator"
out="$here/_ref"
if [ ! -d "$ref" ]; then
  echo "build_ref: reference sources not present; skipping" >&2
  exit 0
fi
mkdir -p "$out"
cxx_units=(006_get_dependencies_GLU.cpp 007_convert_dependencies_in_levels.cpp
  008_convert_chains_columns_levels_into_vector.cpp
  009_get_max_numbers_per_column.cpp GLU_CSC_CPU_levels_strings.cpp
  GLU_CSC_CPU_Triangular.cpp)

build_variant() {
  local name="$1"; shift
  local flags=("$@")
  local obj="$out/obj_$name"
  mkdir -p "$obj"
  local objs=()
  for f in "$ref"/AmdLib/*.c; do
    local o="$obj/$(basename "${f%.c}").o"
    gcc -fPIC "${flags[@]}" -I"$ref/AmdLib" -c "$f" -o "$o"
    objs+=("$o")
  done
  for u in "${cxx_units[@]}"; do
    local o="$obj/${u%.cpp}.o"
    g++ -std=c++20 -fPIC "${flags[@]}" -I"$ref" -I"$ref/AmdLib" -c "$ref/$u" -o "$o"
    objs+=("$o")
  done
  g++ -std=c++20 -fPIC "${flags[@]}" -I"$ref" -I"$ref/AmdLib" \
    -c "$here/ref_harness.cpp" -o "$obj/ref_harness.o"
  objs+=("$obj/ref_harness.o")
  g++ -shared -o "$out/libpvref$name.so" "${objs[@]}"
  rm -rf "$obj"
}

build_variant "" -O3 -ffp-contract=off
build_variant "_fast" -O3 -march=x86-64-v3 -funroll-loops --fast-math
build_variant "_o3" -O3 -march=x86-64-v3 -funroll-loops
echo "build_ref: $out/libpvref.so $out/libpvref_fast.so $out/libpvref_o3.so"
