#!/bin/bash
# Build the node-code dump harness (oracle/mask_ref_main.c) against the reference's
# own solid_init.c, slip_ridges_init.c, slip_posts_init.c, slip_spanwise_init.c and
# globals.c, read in place, for one parameter set.  Used here only, to write the
# mask fixtures (tests/golden/make_masks.py); nothing on the GPU side runs it.
#
# usage: oracle/build_mask_ref.sh <name> key=value...   (keys as oracle/build_ref.sh)
# output: oracle/_ref/<name>/mask-dump
set -euo pipefail

REF_ROOT=${DNSG_REFERENCE:-/root/reference}
SRC="$REF_ROOT/DNS-solver-with-grid-embedding/src"
HERE=$(cd "$(dirname "$0")" && pwd)
name=$1; shift
OUT="$HERE/_ref/$name"

declare -A P=( [Rebs]=7200.000 [ub]=0.0866034 [xdv]=64 [ydv]=32 [zdv]=32 [TSTR]=0 [TEND]=1
               [CFL]=0.1 [gratio]=2 [zds]=8 [gyplussize]=56 [wyplussize]=8 [gxplussize]=4
               [wxplussize]=4 [pattern]=0 [XDIM]=1 [YDIM]=1 )
for kv in "$@"; do P[${kv%%=*}]=${kv#*=}; done

[ -d "$SRC" ] || { echo "reference sources not found at $SRC" >&2; exit 3; }
export PATH=/opt/conda/bin:$PATH
export MPICH_CC=gcc
command -v mpicc >/dev/null || { echo "mpicc not available" >&2; exit 3; }

TMP=$(mktemp -d /tmp/dnsg_maskbuild.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

sedargs=()
for k in Rebs ub zdv ydv xdv TSTR TEND CFL gratio zds gyplussize wyplussize gxplussize wxplussize XDIM YDIM; do
  sedargs+=(-e "s|^#define $k .*|#define $k ${P[$k]}|")
done
sedargs+=(-e 's|^#define PSLIP|//#define PSLIP|' -e 's|^#define RSLIP|//#define RSLIP|' -e 's|^#define SSLIP|//#define SSLIP|')
case ${P[pattern]} in
  1) sedargs+=(-e 's|^//#define RSLIP|#define RSLIP|') ;;
  2) sedargs+=(-e 's|^//#define PSLIP|#define PSLIP|') ;;
  3) sedargs+=(-e 's|^//#define SSLIP|#define SSLIP|') ;;
esac
sed "${sedargs[@]}" "$SRC/definitions.h" > "$TMP/definitions.h"

objs=()
for b in globals solid_init slip_ridges_init slip_posts_init slip_spanwise_init; do
  (cd "$TMP" && mpicc -O2 -fcommon -w -I"$SRC" -x c -c - -o "$b.o" < "$SRC/$b.c")
  objs+=("$TMP/$b.o")
done
(cd "$TMP" && mpicc -O2 -fcommon -w -I"$SRC" -x c -c - -o main.o < "$HERE/mask_ref_main.c")
mkdir -p "$OUT"
mpicc -o "$OUT/mask-dump" "$TMP/main.o" "${objs[@]}" -lm
for k in "${!P[@]}"; do echo "$k=${P[$k]}"; done | sort > "$OUT/params.txt"
echo "built $OUT/mask-dump"
