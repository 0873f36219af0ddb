"""Queue-driven BFS on CSR graphs (host mirror of the reference's
concurrent_bfs harness, bfs_bench.cpp): synthetic graphs, single BFS runs
and the batched/served paths, all through libgcq's C ABI."""
import ctypes as C
import os

import numpy as np

from . import GLFQ, GWFQ, SFQ, WFQ, _check, load  # noqa: F401
from .workloads import BatchConfig, BatchResult, BfsConfig, BfsResult, QueryResult, TuneTrial

HOST_LOOP, PERSISTENT, ASYNC, BITMAP = 0, 1, 2, 3


class Graph:
    """An int32 CSR graph; the library uploads it to HBM on first use."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def lattice(cls, rows, cols, keep=0.75, diag=0.01, seed=7):
        h = C.c_void_p()
        _check(load().gcq_graph_lattice(rows, cols, keep, diag, seed, C.byref(h)), "gcq_graph_lattice")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=11):
        h = C.c_void_p()
        _check(load().gcq_graph_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)), "gcq_graph_rmat")
        return cls(h)

    @classmethod
    def load(cls, path, fmt="auto", undirected=True):
        """A Matrix Market (.mtx) or SNAP edge-list file, read with the
        reference loaders' rules (bfs_bench.cpp:202-243); fmt "auto" picks by
        the ".mtx" suffix."""
        code = {"auto": 0, "mtx": 1, "snap": 2}[fmt]
        h = C.c_void_p()
        _check(load().gcq_graph_load(os.fsencode(path), code, int(bool(undirected)), C.byref(h)), "gcq_graph_load")
        return cls(h)

    @classmethod
    def from_csr(cls, offsets, indices):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        h = C.c_void_p()
        _check(load().gcq_graph_from_csr(len(off) - 1, len(idx), off.ctypes.data, idx.ctypes.data if len(idx) else None,
                                         C.byref(h)), "gcq_graph_from_csr")
        return cls(h)

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("graph destroyed")
        return self._h

    def info(self):
        nv, ne = C.c_int32(), C.c_int64()
        _check(load().gcq_graph_info(self.handle, C.byref(nv), C.byref(ne)), "gcq_graph_info")
        return nv.value, ne.value

    def csr(self):
        nv, ne = self.info()
        off = np.empty(nv + 1, dtype=np.int32)
        idx = np.empty(ne, dtype=np.int32)
        _check(load().gcq_graph_copy_csr(self.handle, off.ctypes.data, idx.ctypes.data), "gcq_graph_copy_csr")
        return off, idx

    def pick_source(self, seed, k):
        return int(load().gcq_graph_pick_source(self.handle, seed, k))

    def bfs(self, src, kind=GLFQ, threads=8192, block=256, mode=PERSISTENT, depths=True):
        """One BFS; returns (depth array or None, result dict)."""
        nv, _ = self.info()
        cfg = BfsConfig(kind, threads, block, mode)
        res = BfsResult()
        depth = np.empty(nv, dtype=np.int32) if depths else None
        _check(load().gcq_bfs_run(self.handle, src, C.byref(cfg), depth.ctypes.data if depths else None,
                                  C.byref(res)), "gcq_bfs_run")
        return depth, {f: getattr(res, f) for f, _ in BfsResult._fields_ if f != "pad"}

    def close(self):
        if self._h is not None:
            load().gcq_graph_destroy(self._h)
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


def batch_config(**kw):
    """The library's default batch configuration with keyword overrides."""
    c = BatchConfig()
    load().gcq_batch_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _cfg_dict(c):
    return {f: getattr(c, f) for f, _ in BatchConfig._fields_ if f != "pad"}


class Batch:
    """Batched BFS over resident graphs: many queries in one persistent
    kernel through one shared frontier queue."""

    def __init__(self, graphs):
        self.graphs = list(graphs)  # keep the handles alive
        arr = (C.c_void_p * len(self.graphs))(*[g.handle for g in self.graphs])
        h = C.c_void_p()
        _check(load().gcq_batch_create(arr, len(self.graphs), C.byref(h)), "gcq_batch_create")
        self._h = h

    def run(self, graph_ids, sources, cfg=None, depths=False, **kw):
        """Returns (list of depth arrays or None, per-query dicts, batch dict)."""
        gids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        srcs = np.ascontiguousarray(sources, dtype=np.int32)
        nq = len(gids)
        cfg = cfg if cfg is not None else batch_config(**kw)
        outs, ptrs = None, None
        if depths:
            outs = [np.empty(self.graphs[g].info()[0], dtype=np.int32) for g in gids]
            ptrs = (C.c_void_p * nq)(*[o.ctypes.data for o in outs])
        pq = (QueryResult * nq)()
        res = BatchResult()
        _check(load().gcq_batch_run(self._h, C.byref(cfg), nq, gids.ctypes.data, srcs.ctypes.data, ptrs, pq,
                                    C.byref(res)), "gcq_batch_run")
        per = [{f: getattr(q, f) for f, _ in QueryResult._fields_} for q in pq]
        return outs, per, {f: getattr(res, f) for f, _ in BatchResult._fields_}

    def autotune(self, graph_ids, sources, kind=GLFQ, reps=3, max_trials=256):
        gids = np.ascontiguousarray(graph_ids, dtype=np.int32)
        srcs = np.ascontiguousarray(sources, dtype=np.int32)
        best = BatchConfig()
        trials = (TuneTrial * max_trials)()
        nt = C.c_int32(0)
        _check(load().gcq_batch_autotune(self._h, kind, len(gids), gids.ctypes.data, srcs.ctypes.data, reps,
                                         C.byref(best), trials, max_trials, C.byref(nt)), "gcq_batch_autotune")
        rows = [dict(_cfg_dict(t.cfg), ms=t.ms) for t in trials[: min(nt.value, max_trials)]]
        return best, rows

    def close(self):
        if self._h is not None:
            load().gcq_batch_destroy(self._h)
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


def lattice(rows, cols, keep=0.75, diag=0.01, seed=7):
    """Host CSR (offsets, indices) of the synthetic lattice."""
    with Graph.lattice(rows, cols, keep, diag, seed) as g:
        return g.csr()


def bfs_depths(offsets, indices, src, kind=GLFQ, threads=1024, block=256, mode=PERSISTENT):
    with Graph.from_csr(offsets, indices) as g:
        d, _ = g.bfs(src, kind=kind, threads=threads, block=block, mode=mode)
        return d


def rmat_query_set(nq=256, scale_lo=12, scale_hi=20, edge_factor=16, seed=11):
    """The batched-BFS workload: R-MAT (a, b, c) = (.57, .19, .19) graphs of
    scales scale_lo..scale_hi (one resident graph per scale), nq queries whose
    scale is drawn uniformly (so V is log-uniform) with a random source of
    nonzero degree.  Returns (graphs, graph_ids, sources)."""
    scales = list(range(scale_lo, scale_hi + 1))
    graphs = [Graph.rmat(s, edge_factor, 0.57, 0.19, 0.19, seed) for s in scales]
    rng = np.random.default_rng(seed)
    gids = rng.integers(0, len(scales), nq).astype(np.int32)
    srcs = np.array([graphs[g].pick_source(seed, i) for i, g in enumerate(gids)], dtype=np.int32)
    return graphs, gids, srcs
