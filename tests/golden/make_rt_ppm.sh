#!/usr/bin/env bash
# Writes the golden ray-traced images: the reference's own ray tracing
# harness and G-LFQ queue (concurrent_rt/rt_final.cpp), compiled from its
# sources with -ffp-contract=off by oracle/build_ref.sh (oracle/_ref/
# refnc_rt_glfq), renders its two scenes at 1280 x 720, 4 bounces, 1,024
# threads per tile on one GPU, and its 8-bit PPMs are kept gzip-compressed:
#   rt_glfq_complex.ppm.gz  (scene 0)    rt_glfq_cornell.ppm.gz  (scene 1)
# Usage: bash tests/golden/make_rt_ppm.sh OUTDIR   (needs the GPU and oracle/_ref)
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
OUT=${1:?output directory}
BIN=$ROOT/oracle/_ref/refnc_rt_glfq
[ -x "$BIN" ] || { echo "make_rt_ppm: $BIN is not built (oracle/build_ref.sh)"; exit 1; }
mkdir -p "$OUT"
OUT=$(cd "$OUT" && pwd)
WORK=$(mktemp -d)
cd "$WORK"
for scene in 0 1; do
  echo "[make_rt_ppm] scene $scene"
  timeout -k 10 60 "$BIN" "$scene" 4 1024 > "$OUT/refnc_rt_glfq_$scene.log" 2>&1
done
for name in complex cornell; do gzip -9 -n -c "rt_glfq_$name.ppm" > "$OUT/rt_glfq_$name.ppm.gz"; done
ls -la "$OUT"
