#!/bin/bash
# Builds oracle/_ref/ref_driver from the reference's host-only helper sources
# where they lie under /root/reference (read-only; nothing is copied) plus the
# driver oracle/ref_driver.cpp.  Test infrastructure only: the product never
# links it.  Exits 0 without building when the reference tree is absent.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${REFERENCE_ROOT:-/root/reference}/cpp-src"
if [ ! -d "$REF" ]; then
  echo "reference tree not found at $REF; oracle/_ref not built"
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++17 -O2 -o "$HERE/_ref/ref_driver" \
  -I"$REF/include" -I"$REF/Utilities" -I"$REF/System" \
  "$HERE/ref_driver.cpp" \
  "$REF/Utilities/BSpline.cpp" "$REF/Utilities/BSpmod.cpp" \
  "$REF/Utilities/Scattering.cpp" "$REF/System/Cell.cpp" "$REF/System/AtomCounter.cpp"
echo "built $HERE/_ref/ref_driver"
