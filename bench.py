#!/usr/bin/env python3
"""Headline benchmark: split-role queue throughput on one MI355X.

Workload (the throughput harness configuration): G-WFQ, 65,536 threads at
block 256, 25% producers / 75% consumers (producer iff tid*100/T < 25), 128
operations per thread per launch, ring capacity 2^20, u64 payload
(producer_id << 32) | (seq + 1).  One step = one chunk launch of that
harness; --warmup W untimed launches, then exactly --steps K timed launches
bracketed by a device synchronise.  value = successful enqueues + dequeues per
second over the K launches (Mops/s).

Beside it, measured in the same run:
  * roofline: the launch's successful ops/s against the single-counter
    wave-aggregated FAA peak P1 (tickets/s on one 64-bit counter, same grid);
  * roofline_hbm: algorithmic bytes per launch / mean launch time vs 8 TB/s;
  * cpu_baseline: the oracle's host std::atomic G-WFQ on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one successful operation through the G-WFQ
# indirection: a free-index ticket slot read + consume (8 + 8), the payload
# (8), the published-index slot read + CAS (8 + 8).
BYTES_PER_OP = 40


def _profile_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(kind, threads, run_ms):
    from oracle import pyoracle as orc
    r = orc.tb_run(kind, threads=threads, mode=1, producer_percent=25, chunk_ops=128, warmup_ms=200,
                   run_ms=run_ms, capacity=1 << 20)
    return {"value": round(r["succ_mops"], 3), "unit": "Mops/s", "cores": threads, "kind": "port",
            "sample": f"host std::atomic G-WFQ restatement, split 25/75, {threads} threads, chunk 128, "
                      f"200 ms warmup + {run_ms} ms timed, capacity 2^20 (hardware_concurrency="
                      f"{r['hw_threads']})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120, help="timed chunk launches")
    ap.add_argument("--warmup", type=int, default=25, help="untimed chunk launches")
    ap.add_argument("--threads", type=int, default=65536)
    ap.add_argument("--queue", default="gwfq", choices=["gwfq", "glfq", "wfq", "sfq"])
    ap.add_argument("--cpu-baseline-ms", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the queue path is single-GPU; see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import gcq_amd as g
    kind = g.KINDS[args.queue]
    info = g.device_info()
    T = args.threads

    # Atomic roofline denominators, same grid shape as the harness.
    p1 = g.faa_peak(threads=T, block=256, counters=1, stride_bytes=128, aggregate=True, iters=64, launches=10)
    p4 = g.faa_peak(threads=T, block=256, counters=4, stride_bytes=128, aggregate=True, iters=64, launches=10)

    with g.Queue(kind, capacity=1 << 20, max_threads=max(T, 64)) as q:
        r = q.tb_run(g.SPLIT, threads=T, block=256, chunk_ops=128, producer_percent=25,
                     warmup_launches=args.warmup, timed_launches=args.steps)
    succ = r["enq_success"] + r["deq_success"]
    value = r["succ_mops"]
    ms_per_step = r["elapsed_ms"] / args.steps
    succ_per_launch = succ / args.steps
    kernel_ms = r["kernel_ms_mean"]
    achieved_ops = succ_per_launch / (kernel_ms * 1e-3)

    line = {
        "metric": "split-role queue Mops/s",
        "value": round(value, 3),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"split-role throughput, {g.KIND_NAMES[kind]}, {T} threads, block 256, "
                               "25% producers, 128 ops/thread/launch, capacity 2^20",
                   "queue": args.queue, "threads": T, "block": 256, "chunk_ops": 128, "producer_percent": 25,
                   "capacity": 1 << 20, "parallelism": "single GPU"},
        "roofline": {"bound": "atomic", "achieved": round(achieved_ops / 1e9, 4),
                     "peak": round(p1["tickets_per_s"] / 1e9, 4), "unit": "Gop/s",
                     "frac": round(achieved_ops / p1["tickets_per_s"], 4),
                     "traffic": _profile_traffic("tb_split_gwfq_65536.json"),
                     "peak_kind": "single-counter wave-aggregated FAA (tickets/s), measured in this run",
                     "peak_striped4_gops": round(p4["tickets_per_s"] / 1e9, 3)},
        "roofline_hbm": {"bound": "hbm", "achieved": round(achieved_ops * BYTES_PER_OP / 1e9, 3),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved_ops * BYTES_PER_OP / 1e9 / HBM_PEAK_GBS, 5),
                         "bytes_per_op": BYTES_PER_OP},
        "detail": {"enq_success": r["enq_success"], "deq_success": r["deq_success"], "deq_empty": r["deq_empty"],
                   "enq_fail": r["enq_fail"], "kernel_ms_mean": round(kernel_ms, 4),
                   "elapsed_ms": round(r["elapsed_ms"], 3), "device": info["name"], "cus": info["cus"]},
    }
    if not args.no_cpu_baseline:
        # the host share of one GPU is 16 cores whatever os.cpu_count() reports
        cores = min(len(os.sched_getaffinity(0)), 16)
        line["cpu_baseline"] = cpu_baseline(kind, cores, args.cpu_baseline_ms)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
