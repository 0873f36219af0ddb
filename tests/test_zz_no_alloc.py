"""The allocation-free BFS entry point (gcq_bfs_run_noalloc) through the C
ABI: its depths equal gcq_bfs_run's (G-LFQ, persistent) and the serial
oracle's, bit for bit, on two calls in a row, the second on a smaller graph.
The test allocates the caller's device buffers itself with the HIP runtime
libgcq already loaded; the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_V = 1 << 21


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libgcq.so links."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def to_device(self, a):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        if a.nbytes:
            assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, a):
        assert self.lib.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self.lib.hipFree(p)


def _noalloc(gcq, hip, off, idx, src, threads=8192, block=256):
    from gcq_amd.workloads import BfsResult
    nv = len(off) - 1
    d_off = hip.to_device(np.ascontiguousarray(off, np.int32))
    d_idx = hip.to_device(np.ascontiguousarray(idx, np.int32))
    d_depth = hip.to_device(np.full(nv, 7, np.int32))  # stale contents must be overwritten
    res = BfsResult()
    rc = gcq.load().gcq_bfs_run_noalloc(d_off, d_idx, nv, len(idx), src, threads, block, None, d_depth, C.byref(res))
    assert rc == 0, gcq.load().gcq_last_error().decode()
    got = hip.to_host(d_depth, np.empty(nv, np.int32))
    hip.free(d_off, d_idx, d_depth)
    return got, res


def test_noalloc_bfs_matches_gcq_bfs_run(gcq, orc):
    from gcq_amd import bfs
    hip = _Hip()
    cases = [("rmat 20", bfs.Graph.rmat(20, 16, seed=11), 12345),     # near the 2^21-vertex limit's order, hubs
             ("lattice 300x200", bfs.Graph.lattice(300, 200, keep=0.75, diag=0.01, seed=7), 0)]  # then smaller
    for name, g, src in cases:
        print(f"[noalloc] {name}", flush=True)
        off, idx = g.csr()
        ref, rres = g.bfs(src, kind=gcq.GLFQ, threads=8192)
        got, res = _noalloc(gcq, hip, off, idx, src)
        want, _, _ = orc.bfs_serial(off, idx, src)
        assert np.array_equal(got, ref), name
        assert np.array_equal(got, want), name
        assert res.levels == rres["levels"] and res.reached == rres["reached"]
        assert res.edges_scanned == rres["edges_scanned"]
        g.close()
    print("[noalloc] limit", flush=True)
    lib = gcq.load()
    from gcq_amd.workloads import BfsResult
    r = BfsResult()
    assert lib.gcq_bfs_run_noalloc(1, 1, MAX_V + 1, 0, 0, 256, 256, None, 1, C.byref(r)) == -1
