#!/usr/bin/env bash
# Compiles the reference's PLY loader from its own sources under /root/reference
# (volumetric/src/GaussianGeometry.cpp and its vendored volumetric/src/miniply.cpp, header-only
# torpedo/math) together with ref_ply_capi.cpp into oracle/_ref/libtpd_ref_ply.so.  No reference
# file is copied; nothing is built when /root/reference is absent.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref=/root/reference/torpedo
if [ ! -d "$ref" ]; then echo "ref_ply.sh: $ref absent, skipping"; exit 0; fi
mkdir -p "$here/_ref"
# -include cstdint / cstring: the reference sources use uint32_t and memcpy without including their
# headers (their own toolchain pulled them in transitively; this libstdc++ does not).
/opt/rocm/llvm/bin/clang++ -std=c++23 -O2 -fPIC -shared -ffp-contract=off -include cstdint -include cstring \
  -I"$ref/math/include" -I"$ref/volumetric/include" -I"$ref/volumetric/src" \
  "$here/ref_ply_capi.cpp" "$ref/volumetric/src/GaussianGeometry.cpp" "$ref/volumetric/src/miniply.cpp" \
  -o "$here/_ref/libtpd_ref_ply.so"
echo "built $here/_ref/libtpd_ref_ply.so"
