#!/usr/bin/env bash
# Writes the golden statistics fixtures of this directory: the output line of
# the reference's own CPU oracle (stats_ref, built from the reference sources
# by `make -C oracle ref`) on fixed windows [START, LIMIT]. Each fixture file
# holds one line "START LIMIT <stats_ref output fields>". Run in the
# development container (the reference tree is not on the GPU machines):
#     bash tests/golden/make_fixtures.sh [name ...]
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
ref=$here/../../oracle/_ref/stats_ref
[ -x "$ref" ] || { echo "build the reference oracle first: make -C oracle ref"; exit 1; }
# name start limit
windows=(
  "w1e6        1000000               3000000"
  "w1e12s      1000000000000         1000002000000"
  "w4e18s      4000000000000000000   4000000000002000000"
  "w2p63s      9223372036853775808   9223372036855775808"
  "wgap1476    1425172824437600000   1425172824437800000"
  "wgap1442    804212830686600000    804212830686800000"
  "wtop        18446742974194923840  18446742974197923840"
  "w4e18       4000000000000000000   4000000000020000000"
  "w44e18      4400000000000000000   4400000000010000000"
  "w5e18end    4999999999980000000   5000000000000000000"
  "w1e12       1000000000000         1000010000000"
  "w2p63       9223372036852775808   9223372036857775808"
  "wgapcross   1425172824436700150   1425172824438700150"
  "w4e18l      4000000000000000000   4000000000200000000"
  "cfg1        1000000000000         1000467499998"
  "cfg2        4000000000000000000   4000000000467499998"
)
want=" $* "
for w in "${windows[@]}"; do
  set -- $w
  name=$1 a=$2 b=$3
  if [ "$want" != "  " ] && [[ "$want" != *" $name "* ]]; then continue; fi
  out=$here/ref_$name.txt
  [ -s "$out" ] && continue
  t0=$(date +%s)
  line=$("$ref" "$a" "$b")
  echo "$a $b $line" > "$out.tmp" && mv "$out.tmp" "$out"
  echo "$name [$a, $b]: $(( $(date +%s) - t0 )) s"
done
