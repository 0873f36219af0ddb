"""Queue-driven BFS on CSR graphs (host mirror of the reference's
concurrent_bfs harness, bfs_bench.cpp): synthetic graphs, single BFS runs
and the batched/served paths, all through libgcq's C ABI."""
import ctypes as C

import numpy as np

from . import GLFQ, GWFQ, SFQ, WFQ, _check, load  # noqa: F401
from .workloads import BfsConfig, BfsResult

HOST_LOOP, PERSISTENT = 0, 1


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


def lattice(rows, cols, keep=0.75, diag=0.01, seed=7):
    """Host CSR (offsets, indices) of the synthetic lattice."""
    with Graph.lattice(rows, cols, keep, diag, seed) as g:
        return g.csr()


def bfs_depths(offsets, indices, src, kind=GLFQ, threads=1024, block=256, mode=PERSISTENT):
    with Graph.from_csr(offsets, indices) as g:
        d, _ = g.bfs(src, kind=kind, threads=threads, block=block, mode=mode)
        return d
