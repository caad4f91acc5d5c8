#!/bin/bash
# Builds the reference's own CPU runtime (KunQuant cpp/Kun + the LogFast
# tables) straight from its sources under /root/reference into oracle/_ref/.
# Test infrastructure only: used by tests/golden/gen_ref_vectors.py to write
# golden vectors. Flags follow the reference's JIT (cfake.py:91-110):
# -std=c++11 (strict ISO: no FP contraction) -O3 with AVX2/FMA.
set -e
REF=${KUN_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF/cpp/Kun" ] || { echo "reference sources not found at $REF"; exit 0; }
mkdir -p "$OUT"
SRCS="$REF/cpp/Kun/Aggregration.cpp $REF/cpp/Kun/CApi.cpp $REF/cpp/Kun/CorrWith.cpp $REF/cpp/Kun/Executor.cpp \
      $REF/cpp/Kun/IO.cpp $REF/cpp/Kun/Module.cpp $REF/cpp/Kun/Rank.cpp $REF/cpp/Kun/Runtime.cpp \
      $REF/cpp/Kun/Scale.cpp $REF/cpp/Kun/SkipList.cpp $REF/cpp/KunSIMD/cpu/tables.cpp"
g++ -std=c++11 -O3 -mavx2 -mfma -ffp-contract=off -fPIC -shared -DKUN_CORE_LIB -I "$REF/cpp" $SRCS -o "$OUT/libKunRuntime.so.tmp" -ldl -lpthread
mv "$OUT/libKunRuntime.so.tmp" "$OUT/libKunRuntime.so"
echo "built $OUT/libKunRuntime.so"
