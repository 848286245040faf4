#!/usr/bin/env bash
# Builds oracle/_ref/libsph2d_ref.so from the reference's own Fortran sources,
# where they lie under /root/reference (read-only). Only the four files that
# compile on their own are used (no MPI, no ifort extensions):
#   Taylor_Green_Vortex/information_module.f90   constants (exported as symbols)
#   Taylor_Green_Vortex/function_module.f90      MUSCL limiter, 2x2/3x3 inverse, BRINV
#   Taylor_Green_Vortex/compute_kernel.f90       kernel W, dW/dx (improved Gaussian)
#   Taylor_Green_Vortex/status_equation.f90      Tait EOS
# The rest of the reference needs `use MPI` (absent from this image) and
# `inquire(directory=)` (ifort-only), so it is not built.
# The library is test infrastructure: tests load it to pin the C oracle.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ref="${SPH2D_REFERENCE:-/root/reference}/Taylor_Green_Vortex"
fc="${FC:-/opt/rocm/lib/llvm/bin/flang}"
out="$here/_ref"
if [ ! -d "$ref" ]; then
  echo "reference sources not present ($ref); skipping _ref build" >&2
  exit 0
fi
if [ ! -x "$fc" ]; then
  echo "flang not found at $fc; skipping _ref build" >&2
  exit 0
fi
mkdir -p "$out"
# Default x86-64 target: no FMA instructions, so no contraction (matches
# the oracle's -ffp-contract=off).
"$fc" -O2 -fPIC -shared -J "$out" -o "$out/libsph2d_ref.so" \
  "$ref/information_module.f90" \
  "$ref/function_module.f90" \
  "$ref/compute_kernel.f90" \
  "$ref/status_equation.f90" 2> "$out/build.log" || { cat "$out/build.log" >&2; exit 1; }
echo "built $out/libsph2d_ref.so"
