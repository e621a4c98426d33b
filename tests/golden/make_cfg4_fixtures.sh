#!/usr/bin/env bash
# Writes ref_cfg4_short.txt: the reference's own stats_ref (built by
# `make -C oracle ref`) over the 32 shortest of the 256 batched ranges in
# cfg4_ranges.txt, one line "START LIMIT <stats_ref output fields>" per range
# (as make_fixtures.sh writes them), in the list's order.
#     bash tests/golden/make_cfg4_fixtures.sh [threads]
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
ref=$here/../../oracle/_ref/stats_ref
[ -x "$ref" ] || { echo "build the reference oracle first: make -C oracle ref"; exit 1; }
threads=${1:-8}
out=$here/ref_cfg4_short.txt
grep -v '^#' "$here/cfg4_ranges.txt" | awk '{print NR-1, $1, $2, $2 - $1}' | sort -k4,4n | head -32 | sort -k1,1n |
while read -r i a b len; do
  echo "$a $b $("$ref" "$a" "$b" "$threads")"
done > "$out.tmp"
mv "$out.tmp" "$out"
wc -l "$out"
