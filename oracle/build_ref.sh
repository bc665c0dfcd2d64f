#!/usr/bin/env bash
# Builds the reference Corrfunc theory library (the compiled CPU reference used to
# pin the oracle) from the sources under /root/reference, without running the
# reference's own build system.
#
# What this recipe does, and nothing more:
#   1. copies the C sources of the five theory modes and their utils into a
#      temporary directory outside the repository;
#   2. expands the reference's `.c.src` / `.h.src` templates with the same
#      token substitution its rules.mk applies (DOUBLE -> double or float);
#   3. compiles them with gcc into ONE shared library, oracle/_ref/libcorrfunc_ref.so.
# Nothing is stubbed: every header and source compiled is the reference's own.
# The temporary directory is deleted afterwards; only the .so lands in oracle/_ref/.
#
# Usage: oracle/build_ref.sh [march]      (default march: x86-64-v4)
set -euo pipefail
REF=${CORRFUNC_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
MARCH=${1:-x86-64-v4}
if [ ! -d "$REF/theory" ]; then
    echo "build_ref.sh: reference sources not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/cfref_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$TMP/utils"
cp "$REF"/utils/*.h "$REF"/utils/*.src "$REF"/utils/utils.c "$REF"/utils/progressbar.c \
   "$REF"/utils/cpu_features.c "$TMP/utils/"
expand() {  # $1 = template, $2 = output, $3 = double|float
    local def
    if [ "$3" = double ]; then def="#ifndef DOUBLE_PREC
#define DOUBLE_PREC
#endif"; else def="#ifdef DOUBLE_PREC
#undef DOUBLE_PREC
#endif"; fi
    { echo "$def"; sed -e "/DOUBLE_PREC/!s/DOUBLE/$3/g" "$1"; } > "$2"
}
for t in "$TMP"/utils/*.src; do
    base=$(basename "$t"); stem=${base%.*.src}; ext=${base%.src}; ext=${ext##*.}
    for p in double float; do expand "$t" "$TMP/utils/${stem}_${p}.${ext}" $p; done
done
SRCS=("$TMP/utils/utils.c" "$TMP/utils/progressbar.c" "$TMP/utils/cpu_features.c")
for p in double float; do SRCS+=("$TMP/utils/gridlink_impl_${p}.c" "$TMP/utils/gridlink_utils_${p}.c"); done
for mode in DD:countpairs DDrppi:countpairs_rp_pi DDsmu:countpairs_s_mu xi:countpairs_xi wp:countpairs_wp; do
    dir=${mode%%:*}; lib=${mode##*:}
    mkdir -p "$TMP/$dir"
    cp "$REF/theory/$dir"/*.c "$REF/theory/$dir"/*.h "$REF/theory/$dir"/*.src "$TMP/$dir/"
    for t in "$TMP/$dir"/*.src; do
        base=$(basename "$t"); stem=${base%.*.src}; ext=${base%.src}; ext=${ext##*.}
        for p in double float; do expand "$t" "$TMP/$dir/${stem}_${p}.${ext}" $p; done
    done
    SRCS+=("$TMP/$dir/$lib.c" "$TMP/$dir/${lib}_impl_double.c" "$TMP/$dir/${lib}_impl_float.c")
done
INCS=(-I"$TMP/utils")
for d in DD DDrppi DDsmu xi wp; do INCS+=(-I"$TMP/$d"); done
FLAGS=(-std=c99 -O3 -fPIC -fopenmp -march="$MARCH" -funroll-loops -fno-strict-aliasing
       -DVERSION=\"2.5.3\" -DUSE_OMP -DPERIODIC -DENABLE_MIN_SEP_OPT -DCOPY_PARTICLES
       -D_POSIX_SOURCE=200809L -D_GNU_SOURCE -w)
objs=()
i=0
for s in "${SRCS[@]}"; do
    o="$TMP/obj_$i.o"; i=$((i+1))
    gcc "${FLAGS[@]}" "${INCS[@]}" -I"$(dirname "$s")" -c "$s" -o "$o" &
    objs+=("$o")
done
wait
gcc -shared -fopenmp "${objs[@]}" -lm -o "$OUT/libcorrfunc_ref.so.tmp"
mv "$OUT/libcorrfunc_ref.so.tmp" "$OUT/libcorrfunc_ref.so"
echo "built $OUT/libcorrfunc_ref.so (march=$MARCH)"
