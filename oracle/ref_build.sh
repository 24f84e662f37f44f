#!/bin/bash
# Build and run the reference CPU oracle (mrfarstad/thesis src/stencil_cpu.cu)
# from its own sources for a fixed list of configurations.
# TEST INFRASTRUCTURE ONLY: outputs go to oracle/_ref/ (git-ignored) and
# nothing here is shipped or run on the GPU host.  Needs /root/reference.
#
# Each configuration is one binary because the reference fixes DIM,
# DIMENSIONS, RADIUS and ITERATIONS at compile time (include/constants.h).
# The sources are compiled host-only with hipcc: the reference's .cu files
# mark shared host/device helpers with __host__ __device__, which the HIP
# headers of this image define.
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
[ -d "$REF/src" ] || { echo "reference tree not found at $REF" >&2; exit 1; }
mkdir -p "$OUT/bin" "$OUT/solutions"

# DIM DIMENSIONS RADIUS ITERATIONS
CONFIGS=(
  "8 2 1 1" "8 2 1 2" "8 2 1 3" "16 2 1 10" "16 2 2 10" "16 2 3 7" "16 2 4 10"
  "33 2 1 12" "40 2 8 6" "40 2 16 4" "256 2 1 64"
  "8 3 1 1" "8 3 1 2" "8 3 1 3" "12 3 1 10" "12 3 2 10" "16 3 3 6" "16 3 4 10"
  "20 3 8 4" "64 3 1 100" "48 3 4 20" "40 3 16 3" "36 3 16 2"
)

for cfg in "${CONFIGS[@]}"; do
  read -r dim dims r it <<<"$cfg"
  name="${dim}_${it}_${r}_${dims}d"
  bin="$OUT/bin/stencil_cpu_$name"
  if [ ! -x "$bin" ] || [ "$HERE/ref_harness.cpp" -nt "$bin" ]; then
    hipcc -x hip --offload-host-only -include hip/hip_runtime.h \
      -I "$REF" -I "$REF/include" -O2 \
      -D DIM="$dim" -D DIMENSIONS="$dims" -D RADIUS="$r" -D ITERATIONS="$it" \
      "$HERE/ref_harness.cpp" -o "$bin"
  fi
  (cd "$OUT" && "$bin" > /dev/null)
  [ -f "$OUT/solutions/solution_$name" ] || { echo "missing solution_$name" >&2; exit 1; }
done
echo "reference solutions in $OUT/solutions"
