#!/usr/bin/env bash
# Writes ref_serve_short.txt: the reference's own stats_ref (built by
# `make -C oracle ref`) over the 20 shortest of the 24 served requests
# serve.draw_requests(24, 11) (the serving test's request stream), one line
# "START LIMIT <stats_ref output fields>" per request (as make_fixtures.sh
# writes them), shortest first.
#     bash tests/golden/make_serve_fixtures.sh [threads]
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
ref=$here/../../oracle/_ref/stats_ref
[ -x "$ref" ] || { echo "build the reference oracle first: make -C oracle ref"; exit 1; }
threads=${1:-8}
out=$here/ref_serve_short.txt
python3 -c "
import sys
sys.path.insert(0, '$here/../../goldbach-l0_amd')
import serve
for a, b in sorted(serve.draw_requests(24, 11), key=lambda r: r[1] - r[0])[:20]:
    print(a, b)
" | while read -r a b; do
  echo "$a $b $("$ref" "$a" "$b" "$threads")"
done > "$out.tmp"
mv "$out.tmp" "$out"
wc -l "$out"
