#!/bin/bash
# Build the reference's TKE-budget post-processor (TKE-budget/postprocessing: the
# serial program that averages the per-snapshot budget files over time and writes
# them in wall units) from its own sources, read in place, for one parameter set.
# It is a checker: tests run it on the driver's budget files and on the oracle's,
# and compare what it writes.  Plain C (no MPI), built with gcc.
#
# usage: oracle/build_postproc_ref.sh <name> key=value...
#   keys: Reb ubulk TSTR TEND DT xdv ydv zdv gratio zds
# output: oracle/_ref/<name>/postprocess-budget
#
# The configured definitions.h is written into a temporary directory and every .c
# file is compiled from stdin there, so `#include "definitions.h"` finds it first.
set -euo pipefail

REF_ROOT=${DNSG_REFERENCE:-/root/reference}
SRC="$REF_ROOT/TKE-budget/postprocessing"
HERE=$(cd "$(dirname "$0")" && pwd)
name=$1; shift
OUT="$HERE/_ref/$name"

declare -A P=( [Reb]=7200.0000 [ubulk]=0.0866034 [TSTR]=1130 [TEND]=1134 [DT]=1 [xdv]=512 [ydv]=256
               [zdv]=197 [gratio]=4 [zds]=56 )
for kv in "$@"; do P[${kv%%=*}]=${kv#*=}; done

[ -d "$SRC" ] || { echo "reference sources not found at $SRC" >&2; exit 3; }
command -v gcc >/dev/null || { echo "gcc not available" >&2; exit 3; }

TMP=$(mktemp -d /tmp/dnsg_postbuild.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

sedargs=()
for k in Reb ubulk TSTR TEND DT xdv ydv zdv gratio zds; do
  sedargs+=(-e "s|^#define $k .*|#define $k ${P[$k]}|")
done
sed "${sedargs[@]}" "$SRC/definitions.h" > "$TMP/definitions.h"

objs=()
for f in "$SRC"/*.c; do
  b=$(basename "$f" .c)
  # its file-name buffers hold 20 chars for names of 20 (char fn[20] and
  # sprintf of "turb-fld-2d.%.3d.%.4d", time_stat.c): built without the fortified
  # sprintf check, as the tool builds with its own Makefile (mpicc -O3)
  (cd "$TMP" && gcc -O2 -U_FORTIFY_SOURCE -D_FORTIFY_SOURCE=0 -fno-stack-protector -fcommon -w -I"$SRC" -x c -c - -o "$b.o" < "$f")
  objs+=("$TMP/$b.o")
done
mkdir -p "$OUT"
gcc -o "$OUT/postprocess-budget" "${objs[@]}" -lm
for k in "${!P[@]}"; do echo "$k=${P[$k]}"; done | sort > "$OUT/params.txt"
echo "built $OUT/postprocess-budget"
