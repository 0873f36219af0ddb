"""Small command-line entry points that run one workload, for profiling
under rocprofv3 (kernel traces, HBM byte counters):

    python3 -m gcq_amd.probe copy --bytes 1073741824 --iters 5
    python3 -m gcq_amd.probe tb --launches 20
    python3 -m gcq_amd.probe bfs --runs 3
    python3 -m gcq_amd.probe batch --runs 3
    python3 -m gcq_amd.probe rt --runs 3
    python3 -m gcq_amd.probe serve --runs 1
"""
import argparse
import json
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(prog="gcq_amd.probe")
    ap.add_argument("what", choices=["copy", "tb", "bfs", "batch", "rt", "rtc", "serve"])
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--queue", default="gwfq", choices=["gwfq", "glfq", "wfq", "sfq"])
    a = ap.parse_args(argv)
    import gcq_amd as g
    out = {"what": a.what}
    if a.what == "copy":
        gbps, ms = g.hbm_copy(a.bytes, a.iters)
        out.update(bytes=a.bytes, gbps=gbps, kernel_ms=ms)
    elif a.what == "tb":
        # the headline configuration: split 25/75, 65,536 threads, 128 ops, capacity 2^20
        with g.Queue(g.KINDS[a.queue], capacity=1 << 20, max_threads=65536, poll_backoff=0) as q:
            r = q.tb_run(g.SPLIT, threads=65536, block=256, chunk_ops=128, producer_percent=25,
                         warmup_launches=5, timed_launches=a.launches)
        out.update(r)
    elif a.what == "bfs":
        from gcq_amd import bfs
        with bfs.Graph.lattice(4900, 4900, 0.75, 0.01, 7) as gr:
            for _ in range(a.runs):
                _, res = gr.bfs(0, kind=g.GLFQ, threads=8192, block=256, mode=bfs.PERSISTENT, depths=False)
        out.update(res)
    elif a.what == "batch":
        from gcq_amd import bfs
        graphs, gids, srcs = bfs.rmat_query_set()
        with bfs.Batch(graphs) as b:
            for _ in range(a.runs):
                _, _, res = b.run(gids, srcs, threads=262144, block=512, use_graph=0)
        out.update(res)
        for gr in graphs:
            gr.close()
    elif a.what in ("rt", "rtc"):
        from gcq_amd import rt
        with rt.Scene.preset(rt.RANDOM, 64, 3) as s:
            for _ in range(a.runs):
                if a.what == "rt":
                    _, res = s.render(width=1920, height=1080, tile_w=256, tile_h=256, threads_per_tile=8192, block=512)
                else:
                    _, res = s.render_compact(width=1920, height=1080, block=1024)
        out.update(res)
    elif a.what == "serve":
        # the served line's burst: 300 requests at once, 90% BFS on R-MAT
        # scales 14..22, 10% 256x256 tiles of the 64-sphere 1080p frame
        from gcq_amd import bfs, rt, serve
        graphs = [bfs.Graph.rmat(sc, 16, 0.57, 0.19, 0.19, 11) for sc in range(14, 23)]
        nvs = [gr.info()[0] for gr in graphs]
        rows = serve.poisson_trace(300, 1e9, nvs, (1920, 1080), rt_fraction=0.1, seed=6,
                                   graph_sources=lambda gi, i: graphs[gi].pick_source(11, i))
        with rt.Scene.preset(rt.RANDOM, 64, 3) as scene:
            for _ in range(a.runs):
                with serve.Server(graphs, scene, (1920, 1080), bfs_threads=229376, block=512, rt_threads=32768,
                                  frontier_log2=28) as srv:
                    res = serve.latency_summary(srv.run_trace(rows))
        out.update(res)
        for gr in graphs:
            gr.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
