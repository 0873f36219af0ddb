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

Unless --headline-only, the line also carries the other workloads of the
metric, each measured in this run ("secondary"):
  * queue_sweep: all four queues, split 25/75 and balanced, 1,024-65,536
    threads, plus the FIFO self-check at 65,536 threads x 32 operations;
  * hbm_copy: measured streaming-copy bandwidth (the measured HBM roofline);
  * bfs_lattice: queue-driven BFS on the 4900 x 4900 lattice (keep 0.75, 1%
    diagonals, seed 7, source 0, 8,192 threads, G-LFQ, one launch), 3 warmup
    + 10 timed runs: median MTEPS and algorithmic GB/s (28 B per reached
    vertex + 8 B per scanned edge) against HBM, beside the serial CPU BFS;
  * batched_bfs: 256 R-MAT queries (scales 12..20, seed 11) in one batched
    launch (replayed hipGraph) against 256 plain one-launch-per-query runs;
  * served: a Poisson trace (90% BFS on R-MAT scales 14..22, 10% 256x256
    ray-traced tiles of the 64-sphere scene, 64 GB budget) through the
    continuously batched server and through plain per-request launches:
    p50 / p99 latency and throughput, plus both capacities on a burst;
  * rt: the 1920 x 1080 64-sphere frame, 4 bounces, 256 x 256 tiles: Mrays/s,
    beside the stream-compaction renderer on the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The OpenMP BFS baseline binds its threads close together (set before the
# oracle library first starts OpenMP).
os.environ.setdefault("OMP_PROC_BIND", "close")

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one successful operation through the G-WFQ
# indirection: a free-index ticket slot read + consume (8 + 8), the payload
# (8), the published-index slot read + CAS (8 + 8).
BYTES_PER_OP = 40
# FP32 operations of one traced ray of the 64-sphere scene (contraction off,
# so each is one VALU operation): 17 per sphere test (oc 3, b 5, c 7, disc 2)
# x 64, 13 for the ground plane, about 80 for the hit point, normal, shading
# and reflected ray.  A primary ray starts at the camera, whose per-object
# terms (oc and c per sphere, the plane's offset) the scene computes once: 7
# per sphere test and 7 for the plane remain.  MI355X's published FP32
# vector peak, 157.3 TFLOP/s, counts packed FMAs; plain adds and multiplies
# issue at a quarter of it.
RT_FLOP_PER_RAY = 17 * 64 + 13 + 80
RT_FLOP_PER_PRIMARY = 7 * 64 + 7 + 80
FP32_PEAK_TFLOPS = 157.3


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


QUEUE_NAMES = {0: "G-WFQ", 1: "G-LFQ", 2: "WFQ", 3: "SFQ"}


# The CPU set at start: the OpenMP oracle binds the thread that loads it to
# one core, so the affinity read later can be a single CPU.
_START_MASK = frozenset(os.sched_getaffinity(0))


def host_cores():
    """The host share of one GPU is 16 cores, whatever os.cpu_count() reports."""
    return min(len(_START_MASK), 16)


def cpu_baseline(kind, cores, run_ms, warmup_ms=200):
    """The oracle's host std::atomic restatement of the queue, split 25/75,
    swept over 2, 4, ... host threads up to the host share: value is the
    best rate (cores = the threads that gave it), beside the all-cores rate."""
    from oracle import pyoracle as orc
    # from 2 threads: the split harness makes a lone thread a producer with no consumer
    counts = sorted({1 << i for i in range(1, cores.bit_length()) if (1 << i) <= cores} | {max(cores, 2)})
    sweep = {}
    for n in counts:
        r = orc.tb_run(kind, threads=n, mode=1, producer_percent=25, chunk_ops=128, warmup_ms=warmup_ms,
                       run_ms=run_ms, capacity=1 << 20)
        sweep[n] = r["succ_mops"]
    best = max(sweep, key=sweep.get)
    return {"value": round(sweep[best], 3), "unit": "Mops/s", "cores": best, "kind": "port",
            "all_cores": {"value": round(sweep[max(cores, 2)], 3), "cores": max(cores, 2)},
            "sweep": {str(n): round(v, 3) for n, v in sweep.items()},
            "sample": f"host std::atomic {QUEUE_NAMES[kind]} restatement, split 25/75, swept over "
                      f"{', '.join(map(str, counts))} threads, chunk 128, {warmup_ms} ms warmup + {run_ms} ms timed "
                      f"each, capacity 2^20 (hardware_concurrency={r['hw_threads']}, host share {cores})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120, help="timed chunk launches")
    ap.add_argument("--warmup", type=int, default=25, help="untimed chunk launches")
    ap.add_argument("--threads", type=int, default=65536)
    ap.add_argument("--queue", default=None, choices=["gwfq", "glfq", "wfq", "sfq"],
                    help="queue kind (default gwfq; --profile bfs: glfq)")
    ap.add_argument("--cpu-baseline-ms", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the BFS / batched / served / RT lines")
    ap.add_argument("--no-solo", action="store_true",
                    help="skip the solo-claims comparison run (profiles of the headline kernel alone)")
    ap.add_argument("--profile", choices=PROFILE_TARGETS,
                    help="run one workload only and print its result, for rocprofv3 kernel traces and HBM counters")
    ap.add_argument("--runs", type=int, default=3, help="--profile: repetitions of the workload")
    ap.add_argument("--bytes", type=int, default=1 << 30, help="--profile copy: bytes copied")
    ap.add_argument("--iters", type=int, default=5, help="--profile copy: launches")
    ap.add_argument("--mode", default="persistent", help="--profile bfs: BFS mode")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(each queue lives in one GPU's memory; several GPUs would each run their own queues)",
              file=sys.stderr)
        sys.exit(2)

    if args.profile:
        print(json.dumps(profile_run(args)), flush=True)
        return
    args.queue = args.queue or "gwfq"

    import gcq_amd as g
    kind = g.KINDS[args.queue]
    info = g.device_info()
    T = args.threads

    # Atomic roofline denominators, same grid shape as the harness.
    p1 = g.faa_peak(threads=T, block=256, counters=1, stride_bytes=128, aggregate=True, iters=64, launches=10)
    p4 = g.faa_peak(threads=T, block=256, counters=4, stride_bytes=128, aggregate=True, iters=64, launches=10)
    peaks = {"striped_gtickets_per_s": {}, "per_lane_1_counter_gtickets_per_s": None}
    for k in (2, 8, 16, 32, 64):
        pk = g.faa_peak(threads=T, block=256, counters=k, stride_bytes=128, aggregate=True, iters=64, launches=5)
        peaks["striped_gtickets_per_s"][str(k)] = round(pk["tickets_per_s"] / 1e9, 3)
    peaks["striped_gtickets_per_s"]["1"] = round(p1["tickets_per_s"] / 1e9, 3)
    peaks["striped_gtickets_per_s"]["4"] = round(p4["tickets_per_s"] / 1e9, 3)
    raw = g.faa_peak(threads=T, block=256, counters=1, stride_bytes=128, aggregate=False, iters=4, launches=3)
    peaks["per_lane_1_counter_gtickets_per_s"] = round(raw["tickets_per_s"] / 1e9, 4)
    # One hot line shared by FAAs and polling loads, as the Tail line is in
    # the split run: 25% of the waves add, 75% only load.  Its total access
    # rate bounds a design in which every consumer attempt reads the line.
    mix = g.faa_peak(threads=T, block=256, counters=1, stride_bytes=128, aggregate=True, iters=64, launches=10,
                     poll_percent=75)
    other = g.faa_peak(threads=T, block=256, counters=1, stride_bytes=128, aggregate=True, iters=64, launches=10,
                       poll_percent=75, poll_other_line=True)
    peaks["line_faa_plus_polling"] = {
        "faa_per_s_M": round(mix["atomics_per_s"] / 1e6, 2), "loads_per_s_M": round(mix["loads_per_s"] / 1e6, 2),
        "accesses_per_s_M": round((mix["atomics_per_s"] + mix["loads_per_s"]) / 1e6, 2),
        "faa_per_s_M_pollers_on_other_line": round(other["atomics_per_s"] / 1e6, 2),
        "faa_per_s_M_alone": round(p1["atomics_per_s"] / 1e6, 2)}
    # latency of one dependent device-scope access (the bound of
    # level-synchronous work such as the lattice BFS)
    peaks["round_trip_ns"] = {m: round(g.round_trip(m), 1) for m in g.ROUND_TRIP_MODES}

    # Consumers poll without sleeping between empty attempts (poll_backoff 0):
    # measured fastest for this split (backoff 1 and 2 are 6% and 9% slower
    # without combining, equal within 1% with it).  Idle consumer waves of a
    # workgroup combine their claims and producer waves draw their Tail
    # tickets together (the library's default); the same run without either
    # (solo_claims) is reported beside it.
    with g.Queue(kind, capacity=1 << 20, max_threads=max(T, 64), poll_backoff=0) as q:
        r = q.tb_run(g.SPLIT, threads=T, block=256, chunk_ops=128, producer_percent=25,
                     warmup_launches=args.warmup, timed_launches=args.steps)
        solo = None if args.no_solo else q.tb_run(g.SPLIT, threads=T, block=256, chunk_ops=128, producer_percent=25,
                                                   warmup_launches=args.warmup, timed_launches=args.steps,
                                                   solo_claims=True)
    succ = r["enq_success"] + r["deq_success"]
    value = r["succ_mops"]
    # HIP events on the queue's stream bracket exactly the K timed launches,
    # which run back to back: their mean is the launch duration.
    ms_per_step = r["elapsed_ms"] / args.steps
    succ_per_launch = succ / args.steps
    kernel_ms = r["kernel_ms_mean"]
    achieved_ops = succ_per_launch / (ms_per_step * 1e-3)

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
                   "capacity": 1 << 20, "poll_backoff": 0, "claims": "combined per workgroup when idle",
                   "producer_tickets": "drawn per workgroup",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "atomic", "achieved": round(achieved_ops / 1e9, 4),
                     "peak": round(p1["tickets_per_s"] / 1e9, 4), "unit": "Gop/s",
                     "frac": round(achieved_ops / p1["tickets_per_s"], 4),
                     "traffic": _profile_traffic("tb_split_gwfq_65536.json"),
                     "traffic_source": "profiles/tb_split_gwfq_65536.json: rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                       "launch of this kernel, scaled by a copy kernel (committed, not measured "
                                       "in this run)",
                     "duration": "mean of the timed launches (HIP events on the launch stream)",
                     "peak_kind": "single-counter wave-aggregated FAA (tickets/s), measured in this run",
                     "peak_striped4_gops": round(p4["tickets_per_s"] / 1e9, 3)},
        "roofline_hbm": {"bound": "hbm", "achieved": round(achieved_ops * BYTES_PER_OP / 1e9, 3),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved_ops * BYTES_PER_OP / 1e9 / HBM_PEAK_GBS, 5),
                         "bytes_per_op": BYTES_PER_OP},
        "atomic_peaks": peaks,
        "detail": {"enq_success": r["enq_success"], "deq_success": r["deq_success"], "deq_empty": r["deq_empty"],
                   "enq_fail": r["enq_fail"], "kernel_ms_mean": round(kernel_ms, 4),
                   "solo_claims_mops": round(solo["succ_mops"], 3) if solo else None,
                   "solo_claims_kernel_ms_mean": round(solo["kernel_ms_mean"], 4) if solo else None,
                   "elapsed_ms": round(r["elapsed_ms"], 3), "device": info["name"], "cus": info["cus"]},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(kind, host_cores(), args.cpu_baseline_ms)
    if not args.headline_only:
        line["secondary"] = secondary(g, not args.no_cpu_baseline)
    print(json.dumps(line), flush=True)


PROFILE_TARGETS = ["copy", "tb", "tb-balanced", "bfs", "batch", "rt", "rtc", "serve"]


def profile_run(a):
    """One workload, --runs times, for rocprofv3 (profiles/ give the commands):
      copy         streaming copy of --bytes, --iters launches (calibrates HBM byte counters)
      tb           the headline: split 25/75, 65,536 threads, 128 ops, capacity 2^20, --queue,
                   --warmup + --steps launches
      tb-balanced  the same grid in balanced mode
      bfs          the 4900 x 4900 lattice, G-LFQ, 8,192 threads, --mode
      batch        256 R-MAT queries (scales 12..20) in one batched launch, multi-source groups
      rt / rtc     the 1080p 64-sphere frame, queue renderer / stream compaction
      serve        the served line's burst: 300 requests at once"""
    import gcq_amd as g
    out = {"what": a.profile}
    if a.profile == "copy":
        gbps, ms = g.hbm_copy(a.bytes, a.iters)
        out.update(bytes=a.bytes, gbps=gbps, kernel_ms=ms)
    elif a.profile in ("tb", "tb-balanced"):
        p1 = g.faa_peak(threads=65536, block=256, counters=1, stride_bytes=128, aggregate=True, iters=64, launches=10)
        with g.Queue(g.KINDS[a.queue or "gwfq"], capacity=1 << 20, max_threads=65536, poll_backoff=0) as q:
            r = q.tb_run(g.SPLIT if a.profile == "tb" else 0, threads=65536, block=256, chunk_ops=128,
                         producer_percent=25, warmup_launches=a.warmup, timed_launches=a.steps)
        ms = r["elapsed_ms"] / a.steps
        out.update(r, ms_per_step=ms, p1_tickets_per_s=p1["tickets_per_s"],
                   frac_p1=(r["enq_success"] + r["deq_success"]) / a.steps / (ms * 1e-3) / p1["tickets_per_s"])
    elif a.profile == "bfs":
        from gcq_amd import bfs
        mode = {"persistent": bfs.PERSISTENT, "host": bfs.HOST_LOOP, "async": bfs.ASYNC, "bitmap": bfs.BITMAP}[a.mode]
        with bfs.Graph.lattice(4900, 4900, 0.75, 0.01, 7) as gr:
            for _ in range(a.runs):
                _, res = gr.bfs(0, kind=g.KINDS[a.queue or "glfq"], threads=8192, block=256, mode=mode, depths=False)
        out.update(res)
    elif a.profile == "batch":
        from gcq_amd import bfs
        graphs, gids, srcs = bfs.rmat_query_set()
        with bfs.Batch(graphs) as b:
            for _ in range(a.runs):
                # the configuration the autotuner picked for the bench line on MI355X
                _, _, res = b.run(gids, srcs, threads=65536, block=128, pop_width=32, flush_at=64, capacity_log2=0,
                                  multi_source=1, use_graph=0)
        out.update(res)
        for gr in graphs:
            gr.close()
    elif a.profile in ("rt", "rtc"):
        from gcq_amd import rt
        with rt.Scene.preset(rt.RANDOM, 64, 3) as sc:
            for _ in range(a.runs):
                if a.profile == "rt":
                    _, res = sc.render(width=1920, height=1080, tile_w=256, tile_h=256, threads_per_tile=8192,
                                       block=512)
                else:
                    _, res = sc.render_compact(width=1920, height=1080, block=1024)
        out.update(res)
    elif a.profile == "serve":
        from gcq_amd import bfs, rt, serve
        graphs = [bfs.Graph.rmat(sc, 16, 0.57, 0.19, 0.19, 11) for sc in range(14, 23)]
        nvs = [gr.info()[0] for gr in graphs]
        rows = serve.poisson_trace(300, 1e9, nvs, (1920, 1080), rt_fraction=0.1, seed=6,
                                   graph_sources=lambda gi, i: graphs[gi].pick_source(11, i))
        with rt.Scene.preset(rt.RANDOM, 64, 3) as scene:
            for _ in range(a.runs):
                with serve.Server(graphs, scene, (1920, 1080), bfs_threads=245760, block=512, rt_threads=16384,
                                  frontier_log2=28, small_log2=18, small_threads=32768, multi_source=1) as srv:
                    res = serve.latency_summary(srv.run_trace(rows))
        out.update(res)
        for gr in graphs:
            gr.close()
    return out


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def secondary(g, with_cpu):
    import numpy as np

    from gcq_amd import bfs, rt, serve
    out = {}
    gbps, copy_ms = g.hbm_copy(1 << 30, 10)
    out["hbm_copy"] = {"gbps": round(gbps, 1), "bytes": 1 << 30, "kernel_ms": round(copy_ms, 4)}
    _progress(f"hbm copy {gbps:.0f} GB/s")

    # --- all four queues: split 25/75 and balanced, threads swept, capacity 2^20
    sweep = {}
    names = {g.GWFQ: "gwfq", g.GLFQ: "glfq", g.WFQ: "wfq", g.SFQ: "sfq"}
    for kind in (g.GWFQ, g.GLFQ, g.WFQ, g.SFQ):
        rows = {}
        # WFQ's cell pool is its default: 2^22 cells at 65,536 threads, a ring
        # reclaimed as the queue runs
        for mode, mname in ((g.SPLIT, "split25"), (0, "balanced")):
            for T in (1024, 2048, 4096, 8192, 16384, 32768, 65536):
                with g.Queue(kind, capacity=1 << 20, max_threads=65536, poll_backoff=0) as q:
                    r = q.tb_run(mode, threads=T, block=256, chunk_ops=128, producer_percent=25,
                                 warmup_launches=5, timed_launches=20)
                rows[f"{mname}_{T}"] = round(r["succ_mops"], 1)
        with g.Queue(kind, capacity=1 << 20, max_threads=65536) as q:
            f = q.fifo_check(65536, 32, concurrent=True)
        rows["fifo_check_65536x32_verified"] = bool(f["verified"])
        if with_cpu:
            rows["cpu_baseline"] = cpu_baseline(kind, host_cores(), 300, warmup_ms=50)
        sweep[names[kind]] = rows
    out["queue_sweep"] = {"unit": "Mops/s (successful enq + deq)", "chunk_ops": 128, "block": 256,
                          "launches": "5 warmup + 20 timed", "capacity": 1 << 20, "queues": sweep}
    _progress("queue sweep " + ", ".join(f"{k} {v['split25_65536']}" for k, v in sweep.items()))

    # --- plain BFS on the lattice: every frontier queue, persistent mode
    # (3 warmup + 10 timed runs each); the line reports the fastest
    with bfs.Graph.lattice(4900, 4900, 0.75, 0.01, 7) as gr:
        nv, ne = gr.info()
        per_q = {}
        for qn in ("glfq", "gwfq", "wfq", "sfq"):
            kind = g.KINDS[qn]
            for _ in range(3):
                gr.bfs(0, kind=kind, threads=8192, block=256, mode=bfs.PERSISTENT, depths=False)
            ms = []
            for _ in range(10):
                _, res = gr.bfs(0, kind=kind, threads=8192, block=256, mode=bfs.PERSISTENT, depths=False)
                ms.append(res["kernel_ms"])
            per_q[qn] = float(np.median(ms))
        best = min(per_q, key=per_q.get)
        med = per_q[best]
        alg_bytes = 28 * res["reached"] + 8 * res["edges_scanned"]
        rt_ns = g.round_trip("load")
        o = {"workload": f"lattice 4900x4900, keep 0.75, diag 0.01, seed 7, src 0, 8192 threads, persistent, "
                         f"{QUEUE_NAMES[g.KINDS[best]]} frontier (fastest of the four)",
             "vertices": nv, "edges": ne, "levels": res["levels"], "reached": res["reached"],
             "edges_scanned": res["edges_scanned"], "queue": best, "kernel_ms_median": round(med, 3),
             "mteps": round(res["edges_scanned"] / med / 1e3, 1),
             "per_queue_ms": {k: round(v, 3) for k, v in per_q.items()},
             # each level is a chain of dependent device-scope accesses: its
             # time in units of one measured round trip
             "latency_bound": {"per_level_us": round(med * 1e3 / res["levels"], 3),
                               "round_trip_ns": round(rt_ns, 1),
                               "round_trips_per_level": round(med * 1e6 / res["levels"] / rt_ns, 2)},
             "roofline": {"bound": "hbm", "achieved": round(alg_bytes / med / 1e6, 2), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(alg_bytes / med / 1e6 / HBM_PEAK_GBS, 5),
                          "bytes_per_launch": alg_bytes}}
        # the same traversal in the other one-launch modes, same queue (1 warmup + 5 timed)
        o["other_modes"] = {}
        for mname, mode in (("bitmap", bfs.BITMAP), ("async", bfs.ASYNC)):
            gr.bfs(0, kind=g.KINDS[best], threads=8192, block=256, mode=mode, depths=False)
            mm = []
            for _ in range(5):
                _, r2 = gr.bfs(0, kind=g.KINDS[best], threads=8192, block=256, mode=mode, depths=False)
                mm.append(r2["kernel_ms"])
            m2 = float(np.median(mm))
            o["other_modes"][mname] = {"kernel_ms_median": round(m2, 3),
                                       "mteps": round(res["edges_scanned"] / m2 / 1e3, 1),
                                       "expansions_over_edges": round(r2["edges_scanned"] / res["edges_scanned"], 3)}
        if with_cpu:
            from oracle import pyoracle as orc
            off, idx = gr.csr()
            t = time.perf_counter()
            _, _, es = orc.bfs_serial(off, idx, 0)
            cpu_s = time.perf_counter() - t
            o["cpu_baseline"] = {"value": round(es / cpu_s / 1e6, 2), "unit": "MTEPS", "cores": 1, "kind": "port",
                                 "sample": "one serial textbook BFS (oracle) on the same CSR"}
            cores = host_cores()
            orc.bfs_parallel(off, idx, 0, threads=cores)  # warm: first-touch of the frontier arrays
            par = []
            for _ in range(3):
                t = time.perf_counter()
                orc.bfs_parallel(off, idx, 0, threads=cores)
                par.append(time.perf_counter() - t)
            par_s = float(np.median(par))
            o["cpu_baseline_parallel"] = {"value": round(es / par_s / 1e6, 2), "unit": "MTEPS", "cores": cores,
                                          "kind": "port",
                                          "sample": "level-synchronous top-down OpenMP BFS (oracle) on the same CSR, "
                                                    "OMP_PROC_BIND=close, median of 3 after one warm run"}
    out["bfs_lattice"] = o
    _progress(f"lattice BFS {o['mteps']} MTEPS")

    # --- batched vs plain (256 R-MAT queries)
    graphs, gids, srcs = bfs.rmat_query_set()
    plain_threads = 32768
    for gr in graphs:
        gr.bfs(0, kind=g.GLFQ, threads=plain_threads, depths=False)
    t = time.perf_counter()
    for gi, sv in zip(gids, srcs):
        graphs[gi].bfs(int(sv), kind=g.GLFQ, threads=plain_threads, depths=False)
    plain_s = time.perf_counter() - t
    with bfs.Batch(graphs) as b:
        # block, grid, pop width, flush width and ring capacity from the
        # autotuner (coordinate descent, one timed run per trial)
        t = time.perf_counter()
        cfg, trials = b.autotune(gids, srcs, kind=g.GLFQ, reps=1, max_trials=64)
        tune_s = time.perf_counter() - t
        cfg.use_graph = 1
        b.run(gids, srcs, cfg=cfg)
        ts = []
        for _ in range(5):
            _, per, res = b.run(gids, srcs, cfg=cfg)
            ts.append(res["total_ms"])
        bms = float(np.median(ts))
    # 28 B per reached vertex + 8 B per scanned edge, summed over the queries
    alg = sum(28 * q["reached"] + 8 * q["edges_scanned"] for q in per)
    out["batched_bfs"] = {
        "workload": "256 R-MAT queries, scales 12..20 uniform, ef 16, seed 11, G-LFQ",
        "queries_per_s": round(256 / bms * 1e3, 1), "batch_ms_median": round(bms, 2),
        "plain_queries_per_s": round(256 / plain_s, 1), "vs_plain": round(plain_s * 1e3 / bms, 2),
        "plain_threads": plain_threads,
        "roofline": {"bound": "hbm", "achieved": round(alg / bms / 1e6, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / bms / 1e6 / HBM_PEAK_GBS, 5), "bytes_per_launch": alg,
                     "bytes_kind": "per-query algorithmic bytes (28 B per reached vertex + 8 B per scanned edge, "
                                   "summed over queries); multi-source groups scan an edge once per level for all "
                                   "their members, so the launch moves fewer (profiles/batch_rmat256.md)"},
        "config": {k: getattr(cfg, k) for k in ("threads", "block", "pop_width", "flush_at", "capacity_log2",
                                                 "use_graph", "multi_source")},
        "autotune": {"trials": len(trials), "seconds": round(tune_s, 1)},
        "replayed_graph": bool(res["graph_replayed"])}
    for gr in graphs:
        gr.close()
    _progress(f"batched {out['batched_bfs']['queries_per_s']} q/s vs plain {out['batched_bfs']['plain_queries_per_s']}")

    # --- served vs plain (Poisson mix)
    graphs = [bfs.Graph.rmat(sc, 16, 0.57, 0.19, 0.19, 11) for sc in range(14, 23)]
    scene = rt.Scene.preset(rt.RANDOM, 64, 3)
    frame = (1920, 1080)
    nvs = [gr.info()[0] for gr in graphs]
    pick = lambda gi, i: graphs[gi].pick_source(11, i)  # noqa: E731
    for gr in graphs:
        gr.bfs(0, kind=g.GLFQ, threads=32768, depths=False)
    # Poisson traces of 600 requests at four rates, up to and past the plain
    # path's capacity, served and plain; then one 300-request burst each.
    rates = (150.0, 300.0, 450.0, 600.0)
    traces = {lam: serve.poisson_trace(600, lam, nvs, frame, rt_fraction=0.1, seed=5, graph_sources=pick)
              for lam in rates}
    burst = serve.poisson_trace(300, 1e9, nvs, frame, rt_fraction=0.1, seed=6, graph_sources=pick)
    # two BFS lanes: graphs of at most 2^18 vertices (scales 14..18) get
    # 32,768 of the BFS threads and their own level barrier; tiles 16,384
    # threads; queries on one graph share bit-parallel groups (a sweep of
    # splits on MI355X: this one has the lowest p50 at every rate and the
    # highest burst capacity)
    scfg = dict(bfs_threads=245760, block=512, rt_threads=16384, frontier_log2=28, small_log2=18,
                small_threads=32768, multi_source=1)
    pcfg = dict(bfs_threads=65536, rt_threads_per_tile=8192)
    by_rate = {}
    for lam in rates:
        with serve.Server(graphs, scene, frame, **scfg) as srv:
            comps = srv.run_trace(traces[lam])
        s_tr = serve.latency_summary(comps)
        p_tr = serve.latency_summary(serve.plain_run_trace(graphs, scene, frame, traces[lam], **pcfg))
        by_rate[lam] = (s_tr, p_tr, comps)
    rows = traces[rates[0]]
    comps = by_rate[rates[0]][2]
    by_lane = {lane: serve.latency_summary([c for c, r in zip(comps, rows) if r["type"] == 0 and
                                           (nvs[r["graph"]] <= 1 << 18) == (lane == "small")])
               for lane in ("small", "big")}
    by_lane["tiles"] = serve.latency_summary([c for c, r in zip(comps, rows) if r["type"] == 1])
    with serve.Server(graphs, scene, frame, **scfg) as srv:
        s_burst = serve.latency_summary(srv.run_trace(burst))
    p_burst = serve.latency_summary(serve.plain_run_trace(graphs, scene, frame, burst, **pcfg))
    # the same lowest-rate trace and burst without groups (per-query items)
    per_q = dict(scfg, multi_source=0)
    with serve.Server(graphs, scene, frame, **per_q) as srv:
        s_perq = serve.latency_summary(srv.run_trace(rows))
    with serve.Server(graphs, scene, frame, **per_q) as srv:
        s_perq_burst = serve.latency_summary(srv.run_trace(burst))
    rnd = lambda d: {k: round(v, 3) if isinstance(v, float) else v for k, v in d.items()}  # noqa: E731
    trip = lambda d: [round(d["p50_ms"], 2), round(d["p99_ms"], 2), round(d["throughput_rps"], 1)]  # noqa: E731
    s_top, p_top, _ = by_rate[rates[-1]]
    out["served"] = {
        "workload": "Poisson arrivals at 150 / 300 / 450 / 600 requests/s, 600 requests each (and a 300-request "
                    "burst): 90% BFS on R-MAT scales 14..22, 10% 256x256 tiles of the 1920x1080 64-sphere scene "
                    "(4 bounces), 64 GB budget, arrival seed 5 (burst seed 6)",
        "served": rnd(by_rate[rates[0]][0]), "plain": rnd(by_rate[rates[0]][1]),
        "by_rate": {f"{lam:.0f}": {"served_p50_p99_rps": trip(by_rate[lam][0]),
                                   "plain_p50_p99_rps": trip(by_rate[lam][1])} for lam in rates},
        "served_by_class": {k: rnd(v) for k, v in by_lane.items()},
        "served_per_query_items": rnd(s_perq), "burst_served_per_query_items": rnd(s_perq_burst),
        "burst_served": rnd(s_burst), "burst_plain": rnd(p_burst),
        "capacity_vs_plain": round(s_burst["throughput_rps"] / p_burst["throughput_rps"], 2),
        "config": dict(scfg, plain_bfs_threads=pcfg["bfs_threads"],
                       plain_rt_threads_per_tile=pcfg["rt_threads_per_tile"])}
    for gr in graphs:
        gr.close()
    _progress(f"served at {rates[-1]:.0f}/s p50 {s_top['p50_ms']:.1f} / p99 {s_top['p99_ms']:.1f} ms vs plain "
              f"{p_top['p50_ms']:.1f} / {p_top['p99_ms']:.1f} ms; burst capacity "
              f"{out['served']['capacity_vs_plain']}x plain")

    # --- ray tracing
    rcfg = dict(width=1920, height=1080, tile_w=256, tile_h=256, threads_per_tile=8192, block=512)
    scene.render(**rcfg)
    ms = []
    for _ in range(5):
        _, res = scene.render(**rcfg)
        ms.append(res["kernel_ms"])
    med = float(np.median(ms))
    out["rt"] = {"workload": "1920x1080, 64 random spheres + ground, seed 3, 4 bounces, 256x256 tiles, G-LFQ",
                 "rays": res["total_rays"], "kernel_ms_median": round(med, 3),
                 "mrays_per_s": round(res["total_rays"] / med / 1e3, 1),
                 "config": {"threads_per_tile": 8192, "block": 512}}
    n_primary = 1920 * 1080
    rt_flop = RT_FLOP_PER_PRIMARY * n_primary + RT_FLOP_PER_RAY * (res["total_rays"] - n_primary)
    rt_tflops = rt_flop / med / 1e9
    out["rt"]["roofline"] = {"bound": "fp32", "achieved": round(rt_tflops, 2), "peak": FP32_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(rt_tflops / FP32_PEAK_TFLOPS, 4),
                             "flop_per_ray": {"primary": RT_FLOP_PER_PRIMARY, "bounce": RT_FLOP_PER_RAY},
                             "note": "no FMA contraction (bit-exact shading), so plain FP32 adds and multiplies "
                                     "cap it at a quarter of the packed-FMA peak"}
    # the stream-compaction comparator on the same frame (same image)
    scene.render_compact(width=1920, height=1080, block=1024)
    cms = []
    for _ in range(5):
        _, cres = scene.render_compact(width=1920, height=1080, block=1024)
        cms.append(cres["kernel_ms"])
    cmed = float(np.median(cms))
    out["rt"]["compaction"] = {"frame_ms_median": round(cmed, 3), "rays": cres["total_rays"],
                               "mrays_per_s": round(cres["total_rays"] / cmed / 1e3, 1),
                               "queue_frame_over_compaction_frame": round(med / cmed, 3)}
    scene.close()
    return out


if __name__ == "__main__":
    main()
