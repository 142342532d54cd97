#!/bin/bash
# Build the reference solver (DNS-solver-with-grid-embedding/src) from its own
# sources, read in place, for one parameter set.  Used here only, to pin the
# oracle (tests/golden/make_golden.py); nothing on the GPU side runs it.
#
# usage: oracle/build_ref.sh <name> key=value...
#   keys: Rebs ub xdv ydv zdv TSTR TEND CFL gratio zds gyplussize wyplussize
#         gxplussize wxplussize pattern(0..3) ITER XDIM YDIM fltr(0 NOFLTR, 1 FLTR)
#         xdu ydu zdu (the uniform grid read by the '-i' start)
# output: oracle/_ref/<name>/run-dns-mgrd
#
# The reference is configured by editing #define values in definitions.h
# (its README, "Installation" step 2).  The recipe writes that configured header
# into a temporary directory and compiles every reference .c file from stdin in
# that directory, so `#include "definitions.h"` resolves to the configured header
# first and fun_defs.h to the reference's own.  MPI comes from the image's MPICH.
set -euo pipefail

REF_ROOT=${DNSG_REFERENCE:-/root/reference}
SRC="$REF_ROOT/DNS-solver-with-grid-embedding/src"
HERE=$(cd "$(dirname "$0")" && pwd)
name=$1; shift
OUT="$HERE/_ref/$name"

declare -A P=( [Rebs]=7200.000 [ub]=0.0866034 [xdv]=64 [ydv]=32 [zdv]=32 [TSTR]=0 [TEND]=1
               [CFL]=0.1 [gratio]=2 [zds]=8 [gyplussize]=56 [wyplussize]=8 [gxplussize]=4
               [wxplussize]=4 [pattern]=0 [ITER]=0 [XDIM]=1 [YDIM]=1 [fltr]=0
               [xdu]=512 [ydu]=256 [zdu]=221 )
for kv in "$@"; do P[${kv%%=*}]=${kv#*=}; done

[ -d "$SRC" ] || { echo "reference sources not found at $SRC" >&2; exit 3; }
export PATH=/opt/conda/bin:$PATH
export MPICH_CC=gcc
command -v mpicc >/dev/null || { echo "mpicc not available" >&2; exit 3; }

TMP=$(mktemp -d /tmp/dnsg_refbuild.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

sedargs=()
for k in Rebs ub zdv ydv xdv TSTR TEND CFL gratio zds gyplussize wyplussize gxplussize wxplussize XDIM YDIM \
         xdu ydu zdu; do
  sedargs+=(-e "s|^#define $k .*|#define $k ${P[$k]}|")
done
# surface pattern switches (definitions.h:70-72)
sedargs+=(-e 's|^#define PSLIP|//#define PSLIP|' -e 's|^#define RSLIP|//#define RSLIP|' -e 's|^#define SSLIP|//#define SSLIP|')
case ${P[pattern]} in
  1) sedargs+=(-e 's|^//#define RSLIP|#define RSLIP|') ;;
  2) sedargs+=(-e 's|^//#define PSLIP|#define PSLIP|') ;;
  3) sedargs+=(-e 's|^//#define SSLIP|#define SSLIP|') ;;
esac
if [ "${P[ITER]}" != 0 ]; then sedargs+=(-e "s|^//#define ITER .*|#define ITER ${P[ITER]}|"); fi
# fine->coarse transfer switch (definitions.h:87-90)
if [ "${P[fltr]}" = 1 ]; then sedargs+=(-e 's|^#define NOFLTR|//#define NOFLTR|' -e 's|^//#define FLTR$|#define FLTR|'); fi
sed "${sedargs[@]}" "$SRC/definitions.h" > "$TMP/definitions.h"

objs=()
for f in "$SRC"/*.c; do
  b=$(basename "$f" .c)
  (cd "$TMP" && mpicc -O3 -fcommon -w -I"$SRC" -x c -c - -o "$b.o" < "$f")
  objs+=("$TMP/$b.o")
done
mkdir -p "$OUT"
mpicc -o "$OUT/run-dns-mgrd" "${objs[@]}" -lm
for k in "${!P[@]}"; do echo "$k=${P[$k]}"; done | sort > "$OUT/params.txt"
echo "built $OUT/run-dns-mgrd"
