#!/bin/sh
# Regenerates pcg32_vectors.json from the published PCG32 implementation that
# pyarrow vendors (only available in the development container).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
PCGDIR=$(python3 -c 'import pyarrow, os; print(os.path.join(os.path.dirname(pyarrow.__file__), "include/arrow/vendored"))')
OUT=$(mktemp -d)
g++ -O1 -std=c++17 -I"$PCGDIR" "$HERE/make_pcg_vectors.cpp" -o "$OUT/mkpcg"
"$OUT/mkpcg" > "$HERE/pcg32_vectors.json"
rm -rf "$OUT"
