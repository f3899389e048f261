#!/usr/bin/env bash
# Compiles the reference's camera + math path from its own sources under /root/reference
# (rendering/src/Camera.cpp, extension/src/PerspectiveCamera.cpp, header-only torpedo/math)
# together with ref_camera_capi.cpp into oracle/_ref/libtpd_ref_camera.so.  No reference
# file is copied; nothing is built when /root/reference is absent.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref=/root/reference/torpedo
if [ ! -d "$ref" ]; then echo "ref_camera.sh: $ref absent, skipping"; exit 0; fi
mkdir -p "$here/_ref"
# -include cstdint: the reference headers use uint32_t without including <cstdint> (their own
# toolchain pulled it in transitively; this libstdc++ does not).
/opt/rocm/llvm/bin/clang++ -std=c++23 -O2 -fPIC -shared -ffp-contract=off -include cstdint \
  -I"$ref/math/include" -I"$ref/rendering/include" -I"$ref/extension/include" \
  "$here/ref_camera_capi.cpp" "$ref/rendering/src/Camera.cpp" "$ref/extension/src/PerspectiveCamera.cpp" \
  -o "$here/_ref/libtpd_ref_camera.so"
echo "built $here/_ref/libtpd_ref_camera.so"
