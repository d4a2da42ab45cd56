#!/usr/bin/env bash
# Builds oracle/_ref/libqsr_ref.so from the reference's own QSR.cpp (read in
# place under /root/reference, never copied) plus oracle/ref_harness.cpp.
# QSR.cpp's interactive main() is renamed so its free functions can be called.
# Flags follow the reference's documented build line (QSR.h:183:
# g++ -O3 -std=c++23 -fopenmp).  Exits 0 without building when the reference
# tree is absent (a machine without it uses the prebuilt library).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${QSR_REFERENCE_DIR:-/root/reference}"
OUT="$HERE/_ref"
if [[ ! -f "$REF/QSR.cpp" || ! -f "$REF/QSR.h" ]]; then
    echo "build_ref: reference sources not present at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
g++ -O3 -std=c++23 -fopenmp -fPIC -shared -w \
    -Dmain=qsr_reference_interactive_main \
    -I"$REF" "$REF/QSR.cpp" "$HERE/ref_harness.cpp" \
    -o "$OUT/libqsr_ref.so.tmp"
mv "$OUT/libqsr_ref.so.tmp" "$OUT/libqsr_ref.so"
echo "build_ref: wrote $OUT/libqsr_ref.so"
