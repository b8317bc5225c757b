#!/bin/bash
# Builds oracle/_ref/ref_fields{,_fast}: oracle/ref_driver.cpp linked against the
# reference's own field / box / utility sources, compiled where they lie under
# /root/reference (no copies, no stand-ins; Eigen-dependent files are not used).
#   ref_fields       : -O2, IEEE evaluation
#   ref_fields_fast  : the reference's Release flags (-O3 -ffast-math -march=native)
set -euo pipefail
REF=${REF:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
  echo "reference sources not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
if [ -x "$OUT/ref_fields" ] && [ -x "$OUT/ref_fields_fast" ] && [ "$OUT/ref_fields" -nt "$HERE/ref_driver.cpp" ] \
   && [ "$OUT/ref_fields" -nt "$0" ] && [ -z "${FORCE:-}" ]; then
  echo "oracle/_ref is up to date (FORCE=1 rebuilds)"
  exit 0
fi
SRCS="$REF/Fields/MultiPhaseField.cpp $REF/Fields/BaseField.cpp $REF/Boxes/SquareBox.cpp $REF/Boxes/BaseBox.cpp \
  $REF/Interactions/BaseInteraction.cpp $REF/Boxes/SquareWalls.cpp $REF/Boxes/CircleWalls.cpp $REF/Boxes/TriangleWalls.cpp \
  $REF/Utilities/RCexception.cpp $REF/Utilities/Utils.cpp $REF/Utilities/Logger.cpp \
  $REF/Utilities/parse_input/parse_input.cpp $REF/Utilities/parse_input/getline.c"
g++ -std=c++14 -O2 -fopenmp -w -I"$REF" -I"$REF/extern" "$HERE/ref_driver.cpp" $SRCS -o "$OUT/ref_fields"
g++ -std=c++14 -O3 -ffast-math -march=native -fopenmp -w -I"$REF" -I"$REF/extern" "$HERE/ref_driver.cpp" $SRCS -o "$OUT/ref_fields_fast"
echo "built $OUT/ref_fields $OUT/ref_fields_fast"
