#!/usr/bin/env bash
# ORACLE build recipe (test infrastructure only).  Compiles, where the
# reference tree exists, these binaries into oracle/_ref/ (git-ignored; they
# are copied along with the working tree to the machine that runs them):
#   ref_<harness>_<queue>     the reference harness with the reference's own
#                             queue headers, straight from /root/reference;
#   dropin_<harness>_<queue>  the same harness compiled against this
#                             repository's drop-in headers (include/*.hpp);
#   dropinrlx_rt_<queue>      the ray tracing harness on the drop-in headers
#                             without the dequeue acquire (GCQ_DROPIN_RELAXED);
#   refnc_rt_glfq             the reference's ray tracing harness and G-LFQ
#                             queue compiled with -ffp-contract=off, the
#                             arithmetic of this repository's renderer, so its
#                             PPM can be compared byte for byte.
#                             oracle/_ref/dsrc/ mirrors the reference layout
#                             with symbolic links only: the drop-in headers at
#                             its top, links to the harness sources in their
#                             sub-directories, so a harness's "x.hpp" and
#                             "../x.hpp" includes resolve to the drop-ins.  No
#                             reference file is copied.
# Harnesses: tb (throughout_bench.cpp), hist (correctness/history.cpp),
# bfs (concurrent_bfs/bfs_bench.cpp), rt (concurrent_rt/rt_final.cpp).
# Knobs are the reference Makefile's -D macros (Makefile:95-207); the
# reference's own build system is not used.  Incremental: a binary newer
# than its inputs is kept.
set -uo pipefail
REF=${GCQ_REFERENCE:-/root/reference}
[ -d "$REF" ] || { echo "build_ref: no reference tree at $REF, nothing to build"; exit 0; }
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=$ROOT/oracle/_ref
DS=$OUT/dsrc
mkdir -p "$OUT" "$DS/concurrent_bfs" "$DS/concurrent_rt" "$DS/correctness"
for h in gwf_queue.hpp gwf_ring.hpp glf_queue.hpp wfqueue_hip_opt.hpp sfqueue_hip.hpp sfqueue_hip.cpp; do
  ln -sfn "$ROOT/include/$h" "$DS/$h"
done
ln -sfn "$REF/throughout_bench.cpp" "$DS/throughout_bench.cpp"
ln -sfn "$REF/concurrent_bfs/bfs_bench.cpp" "$DS/concurrent_bfs/bfs_bench.cpp"
ln -sfn "$REF/concurrent_rt/rt_final.cpp" "$DS/concurrent_rt/rt_final.cpp"
ln -sfn "$REF/correctness/history.cpp" "$DS/correctness/history.cpp"
ln -sfn "$REF/correctness/wf_linlog.hpp" "$DS/correctness/wf_linlog.hpp"
HIPCC=${HIPCC:-hipcc}
BASE="--offload-arch=gfx950 -O3 -std=c++17 -w -Xclang -target-feature -Xclang -packed-fp32-ops"
TAGS=(gwfq glfq wfq sfq)
DEFS=(USE_GWFQ USE_GLFQ USE_WFQ USE_SFQ)

rel_of() {
  case $1 in
    tb) echo throughout_bench.cpp ;;
    hist) echo correctness/history.cpp ;;
    bfs) echo concurrent_bfs/bfs_bench.cpp ;;
    rt) echo concurrent_rt/rt_final.cpp ;;
  esac
}
knobs_of() {
  case $1 in
    tb) echo "-DONLY_SPLIT=1 -DRUN_MS=400 -DWARMUP_MS=100 -DCHUNK_OPS=128" ;;
    rt) echo "-DRT_MAX_THREADS=1024 -DRT_QUEUE_N=32768" ;;
    *) echo "" ;;
  esac
}

build_one() {  # flavour harness queue-index
  local flav=$1 h=$2 qi=$3 rel src bin log newest=0 t d
  rel=$(rel_of "$h")
  bin=$OUT/${flav}_${h}_${TAGS[$qi]}
  log=$OUT/build_${flav}_${h}_${TAGS[$qi]}.log
  if [ "$flav" = ref ] || [ "$flav" = refnc ]; then src=$REF/$rel; else src=$DS/$rel; fi
  local extra=""
  [ "$flav" = dropinrlx ] && extra="-DGCQ_DROPIN_RELAXED"
  [ "$flav" = refnc ] && extra="-ffp-contract=off"
  for d in "$REF/$rel" "$ROOT"/include/*.hpp "$ROOT"/gcq_amd/csrc/gcq/*.hpp; do
    t=$(stat -L -c %Y "$d"); [ "$t" -gt "$newest" ] && newest=$t
  done
  if [ -x "$bin" ] && [ "$(stat -c %Y "$bin")" -ge "$newest" ]; then return 0; fi
  # shellcheck disable=SC2086
  if ! $HIPCC $BASE -D${DEFS[$qi]} $extra $(knobs_of "$h") -I"$ROOT/gcq_amd/csrc" "$src" -o "$bin" > "$log" 2>&1; then
    echo "build_ref: FAILED $flav $h ${TAGS[$qi]} (oracle/_ref/$(basename "$log"))"
    return 1
  fi
}

items=()
for h in tb hist bfs rt; do for qi in 0 1 2 3; do items+=("dropin $h $qi"); done; done
for qi in 0 1 2 3; do items+=("dropinrlx rt $qi"); done  # drop-in headers without the dequeue acquire
for spec in "tb 0" "tb 1" "tb 2" "tb 3" "hist 0" "hist 1" "hist 2" "hist 3" "bfs 0" "bfs 1" "rt 0" "rt 1"; do
  items+=("ref $spec")
done
items+=("refnc rt 1")
status=0
pids=()
for it in "${items[@]}"; do
  # shellcheck disable=SC2086
  build_one $it &
  pids+=($!)
  while [ "$(jobs -rp | wc -l)" -ge "${GCQ_BUILD_JOBS:-8}" ]; do sleep 0.2; done
done
for p in "${pids[@]}"; do wait "$p" || status=1; done
echo "build_ref: $(ls "$OUT" | grep -cE '^(ref|refnc|dropin|dropinrlx)_[a-z]+_[a-z]+$') binaries in oracle/_ref"
exit $status
