#!/usr/bin/env bash
# Writes ref_cfg3_<index>.txt: the reference's own stats_ref (built by
# `make -C oracle ref`) over each of the 16 segments of the 4.5e18 slice listed
# in cfg3_segments.txt, one line "START LIMIT <stats_ref output fields>" as
# make_fixtures.sh writes. About 6 minutes per segment on 8 threads.
#     bash tests/golden/make_cfg3_fixtures.sh [threads]
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
ref=$here/../../oracle/_ref/stats_ref
[ -x "$ref" ] || { echo "build the reference oracle first: make -C oracle ref"; exit 1; }
threads=${1:-8}
grep -v '^#' "$here/cfg3_segments.txt" | while read -r k a b; do
  out=$here/ref_cfg3_$k.txt
  [ -s "$out" ] && continue
  t0=$(date +%s)
  line=$("$ref" "$a" "$b" "$threads")
  echo "$a $b $line" > "$out.tmp" && mv "$out.tmp" "$out"
  echo "cfg3 segment $k [$a, $b]: $(( $(date +%s) - t0 )) s"
done
