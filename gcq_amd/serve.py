"""Continuously batched serving of BFS queries and ray-traced tiles (one
persistent kernel, admission by slots and a device-memory budget) and the
plain one-launch-per-request comparator, through libgcq's C ABI."""
import ctypes as C

import numpy as np

from . import GLFQ, _check, load  # noqa: F401
from .workloads import BfsConfig, Completion, Request, RtConfig, ServerConfig

BFS, RT = 0, 1


def server_config(**kw):
    c = ServerConfig()
    load().gcq_server_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def requests(rows):
    """rows: dicts with type, graph, src, x0, y0, w, h, keep, id, arrival_ms."""
    arr = (Request * len(rows))()
    for i, r in enumerate(rows):
        for k, v in r.items():
            setattr(arr[i], k, v)
    return arr


def _comps(arr):
    return [{f: getattr(c, f) for f, _ in Completion._fields_ if f != "pad"} for c in arr]


def poisson_trace(n, rate_per_s, graph_nv, frame, tile=256, rt_fraction=0.1, seed=5, keep=0, graph_sources=None):
    """A Poisson arrival trace: (1 - rt_fraction) BFS queries on random
    resident graphs with random sources, rt_fraction tiles of a frame."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1000.0 / rate_per_s, n))
    W, H = frame
    tx, ty = (W + tile - 1) // tile, (H + tile - 1) // tile
    rows = []
    for i in range(n):
        if rng.random() < rt_fraction:
            k = int(rng.integers(0, tx * ty))
            x0, y0 = (k % tx) * tile, (k // tx) * tile
            rows.append(dict(type=RT, x0=x0, y0=y0, w=min(tile, W - x0), h=min(tile, H - y0), keep=keep, id=i,
                             arrival_ms=float(t[i])))
        else:
            g = int(rng.integers(0, len(graph_nv)))
            src = graph_sources(g, i) if graph_sources else int(rng.integers(0, graph_nv[g]))
            rows.append(dict(type=BFS, graph=g, src=src, keep=keep, id=i, arrival_ms=float(t[i])))
    return rows


class Server:
    def __init__(self, graphs, scene=None, frame=(1920, 1080), cfg=None, **kw):
        self.graphs = list(graphs)
        self.scene = scene
        arr = (C.c_void_p * max(1, len(self.graphs)))(*[g.handle for g in self.graphs])
        cfg = cfg if cfg is not None else server_config(**kw)
        h = C.c_void_p()
        _check(load().gcq_server_create(arr, len(self.graphs), scene.handle if scene else None, frame[0], frame[1],
                                        C.byref(cfg), C.byref(h)), "gcq_server_create")
        self._h = h
        self.frame = frame

    def start(self):
        _check(load().gcq_server_start(self._h), "gcq_server_start")

    def submit(self, **r):
        req = Request()
        for k, v in r.items():
            setattr(req, k, v)
        _check(load().gcq_server_submit(self._h, C.byref(req)), "gcq_server_submit")

    def poll(self, max_n=1024):
        out = (Completion * max_n)()
        n = C.c_int32(0)
        _check(load().gcq_server_poll(self._h, out, max_n, C.byref(n)), "gcq_server_poll")
        return _comps(out[: n.value])

    def stop(self):
        _check(load().gcq_server_stop(self._h), "gcq_server_stop")

    def run_trace(self, rows):
        reqs = requests(rows)
        out = (Completion * len(rows))()
        _check(load().gcq_server_run_trace(self._h, reqs, len(rows), out), "gcq_server_run_trace")
        return _comps(out)

    def fetch_depths(self, rid, nv):
        d = np.empty(nv, np.int32)
        _check(load().gcq_server_fetch(self._h, rid, d.ctypes.data), "gcq_server_fetch")
        return d

    def fetch_image(self, rid, w, h):
        img = np.empty((h, w, 3), np.float32)
        _check(load().gcq_server_fetch(self._h, rid, img.ctypes.data), "gcq_server_fetch")
        return img

    def close(self):
        if self._h is not None:
            load().gcq_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass


def plain_run_trace(graphs, scene, frame, rows, bfs_kind=GLFQ, bfs_threads=8192, rt_threads_per_tile=1024):
    """The same trace served one request at a time, one launch per request."""
    arr = (C.c_void_p * max(1, len(graphs)))(*[g.handle for g in graphs])
    b = BfsConfig(bfs_kind, bfs_threads, 256, 1)
    r = RtConfig()
    load().gcq_rt_config_default(C.byref(r))
    r.threads_per_tile = rt_threads_per_tile
    r.tile_w = r.tile_h = 256
    reqs = requests(rows)
    out = (Completion * len(rows))()
    _check(load().gcq_plain_run_trace(arr, len(graphs), scene.handle if scene else None, frame[0], frame[1],
                                      C.byref(b), C.byref(r), reqs, len(rows), out), "gcq_plain_run_trace")
    return _comps(out)


def latency_summary(comps):
    lat = np.array([c["done_ms"] - c["arrival_ms"] for c in comps])
    span = max(c["done_ms"] for c in comps) - min(c["arrival_ms"] for c in comps)
    return dict(n=len(comps), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), throughput_rps=len(comps) / span * 1e3, span_ms=span)
