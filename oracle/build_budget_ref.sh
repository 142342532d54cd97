#!/bin/bash
# Build the reference's TKE-budget computation tool (TKE-budget/computations) from
# its own sources, read in place, for one parameter set.  Used here only, to pin
# the budget oracle (tests/golden/make_golden.py); nothing on the GPU side runs it.
#
# usage: oracle/build_budget_ref.sh <name> key=value...
#   keys: Rebs ub xdv ydv zdv CFL gratio zds TSTR TEND DT
# output: oracle/_ref/<name>/compute-budget
#
# Like oracle/build_ref.sh: the configured definitions.h is written into a
# temporary directory and every tool .c file is compiled from stdin there, so
# `#include "definitions.h"` finds the configured header first.
set -euo pipefail

REF_ROOT=${DNSG_REFERENCE:-/root/reference}
SRC="$REF_ROOT/TKE-budget/computations"
HERE=$(cd "$(dirname "$0")" && pwd)
name=$1; shift
OUT="$HERE/_ref/$name"

declare -A P=( [Rebs]=7200. [ub]=0.0866034 [xdv]=512 [ydv]=256 [zdv]=197 [CFL]=0.1 [gratio]=4
               [zds]=56 [TSTR]=920 [TEND]=1220 [DT]=1 )
for kv in "$@"; do P[${kv%%=*}]=${kv#*=}; done

[ -d "$SRC" ] || { echo "reference sources not found at $SRC" >&2; exit 3; }
export PATH=/opt/conda/bin:$PATH
export MPICH_CC=gcc
command -v mpicc >/dev/null || { echo "mpicc not available" >&2; exit 3; }

TMP=$(mktemp -d /tmp/dnsg_budgetbuild.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

sedargs=()
for k in Rebs ub xdv ydv zdv CFL gratio zds TSTR TEND DT; do
  sedargs+=(-e "s|^#define $k .*|#define $k ${P[$k]}|")
done
sed "${sedargs[@]}" "$SRC/definitions.h" > "$TMP/definitions.h"

objs=()
for f in "$SRC"/*.c; do
  b=$(basename "$f" .c)
  (cd "$TMP" && mpicc -O3 -fcommon -w -I"$SRC" -x c -c - -o "$b.o" < "$f")
  objs+=("$TMP/$b.o")
done
mkdir -p "$OUT"
mpicc -o "$OUT/compute-budget" "${objs[@]}" -lm
for k in "${!P[@]}"; do echo "$k=${P[$k]}"; done | sort > "$OUT/params.txt"
echo "built $OUT/compute-budget"
