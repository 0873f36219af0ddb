"""ORACLE (test infrastructure only): ctypes wrapper around liboracle.so."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OK, ILLEGAL, UNKNOWN = 1, 0, -1
AUTO, WGL, PATTERN = 0, 1, 2
OP_DTYPE = np.dtype([("proc", np.uint32), ("op", np.uint32), ("arg", np.uint64), ("ret", np.uint64),
                     ("call", np.uint64), ("end", np.uint64)])


class HistCfg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("threads", C.c_uint32), ("ops", C.c_uint32), ("mode", C.c_uint32),
                ("producer_percent", C.c_uint32), ("capacity", C.c_uint32), ("enq_patience", C.c_uint32),
                ("deq_patience", C.c_uint32), ("ref_slow_rule", C.c_uint32), ("delay_every", C.c_uint32),
                ("delay_us", C.c_uint32)]


class TbCfg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("threads", C.c_uint32), ("mode", C.c_uint32),
                ("producer_percent", C.c_uint32), ("chunk_ops", C.c_uint32), ("warmup_ms", C.c_uint32),
                ("run_ms", C.c_uint32), ("capacity", C.c_uint32)]


class TbRes(C.Structure):
    _fields_ = [("enq_success", C.c_uint64), ("enq_fail", C.c_uint64), ("deq_success", C.c_uint64),
                ("deq_empty", C.c_uint64), ("elapsed_ms", C.c_double), ("succ_mops", C.c_double),
                ("hw_threads", C.c_uint32)]


_lib = None

# The CPU set this process started with.  liboracle.so is built with OpenMP;
# when OMP_PROC_BIND is set, libgomp binds the thread that loads it (and the
# master of each parallel region) to one place, and std::threads started by
# that thread inherit its mask.  The calls that start std::threads restore the
# full set first.
HOST_MASK = frozenset(os.sched_getaffinity(0))


def _unpin():
    if os.sched_getaffinity(0) != HOST_MASK:
        os.sched_setaffinity(0, HOST_MASK)


def _rt_sig(lib):
    lib.orc_rt_render.restype = C.c_int
    lib.orc_rt_render.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_void_p]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        lib.orc_lincheck.restype = C.c_int
        lib.orc_lincheck.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_char_p, C.c_size_t]
        lib.orc_read_jsonl.restype = C.c_long
        lib.orc_read_jsonl.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t]
        lib.orc_hist_run.restype = C.c_long
        lib.orc_hist_run.argtypes = [C.POINTER(HistCfg), C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32)]
        lib.orc_tb_run.restype = C.c_int
        lib.orc_tb_run.argtypes = [C.POINTER(TbCfg), C.POINTER(TbRes)]
        lib.orc_bfs_serial.restype = C.c_long
        lib.orc_bfs_serial.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                       C.POINTER(C.c_uint64)]
        lib.orc_bfs_parallel.restype = C.c_long
        lib.orc_bfs_parallel.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int]
        _rt_sig(lib)
        _lib = lib
        _unpin()
    return _lib


def to_ops(entries):
    """Convert a history (any structured array with the op fields) to OP_DTYPE."""
    out = np.zeros(len(entries), dtype=OP_DTYPE)
    for f in OP_DTYPE.names:
        out[f] = entries[f]
    return out


def lincheck(entries, engine=AUTO, budget_s=10.0):
    """Return (result, message); result is OK / ILLEGAL / UNKNOWN."""
    ops = np.ascontiguousarray(to_ops(entries))
    msg = C.create_string_buffer(512)
    rc = load().orc_lincheck(ops.ctypes.data, len(ops), engine, budget_s, msg, 512)
    return rc, msg.value.decode()


def read_jsonl(path):
    lib = load()
    n = lib.orc_read_jsonl(path.encode(), None, 0)
    if n < 0:
        raise ValueError(f"cannot parse {path} ({n})")
    out = np.zeros(n, dtype=OP_DTYPE)
    lib.orc_read_jsonl(path.encode(), out.ctypes.data, n)
    return out


def hist_run(kind, threads, ops, mode=0, producer_percent=50, capacity=32768, enq_patience=0, deq_patience=0,
             ref_slow_rule=False, delay_every=0, delay_us=0):
    cfg = HistCfg(kind, threads, ops, mode, producer_percent, capacity, enq_patience, deq_patience,
                  int(ref_slow_rule), delay_every, delay_us)
    cap = threads * ops * 2
    out = np.zeros(cap, dtype=OP_DTYPE)
    steps = C.c_uint32(0)
    lib = load()
    _unpin()
    n = lib.orc_hist_run(C.byref(cfg), out.ctypes.data, cap, C.byref(steps))
    if n < 0:
        raise ValueError("orc_hist_run failed")
    return out[:n], steps.value


def tb_run(kind, threads, mode=1, producer_percent=25, chunk_ops=128, warmup_ms=200, run_ms=1000,
           capacity=1 << 20):
    cfg = TbCfg(kind, threads, mode, producer_percent, chunk_ops, warmup_ms, run_ms, capacity)
    res = TbRes()
    lib = load()
    _unpin()
    if lib.orc_tb_run(C.byref(cfg), C.byref(res)) != 0:
        raise ValueError("orc_tb_run failed")
    return {f: getattr(res, f) for f, _ in TbRes._fields_}


def bfs_serial(offsets, indices, src):
    nv = len(offsets) - 1
    depth = np.empty(nv, dtype=np.int32)
    es = C.c_uint64(0)
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    reached = load().orc_bfs_serial(nv, off.ctypes.data, idx.ctypes.data, src, depth.ctypes.data, C.byref(es))
    return depth, reached, es.value


def bfs_parallel(offsets, indices, src, threads=0):
    nv = len(offsets) - 1
    depth = np.empty(nv, dtype=np.int32)
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    reached = load().orc_bfs_parallel(nv, off.ctypes.data, idx.ctypes.data, src, depth.ctypes.data, threads)
    return depth, reached


def build_rt_variant(out_path, *defines):
    """Compile rt_cpu.cpp alone with extra -D flags (planted-mistake copies for
    tests) and return the loaded library, usable as rt_render(..., lib=...)."""
    flags = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-shared"]
    subprocess.run(["g++", *flags, *[f"-D{d}" for d in defines], os.path.join(HERE, "rt_cpu.cpp"), "-o",
                    str(out_path)], check=True)
    lib = C.CDLL(str(out_path))
    _rt_sig(lib)
    return lib


def rt_render(spheres, planes, cam12, light3, ambient3, W, H, x0=0, y0=0, w=0, h=0, max_depth=4, threads=0,
              lib=None):
    """CPU restatement of the reference ray tracer; returns (image[h, w, 3], rays traced)."""
    w = w or W
    h = h or H
    sph = np.ascontiguousarray(spheres, np.float32).reshape(-1, 8)
    pl = np.ascontiguousarray(planes, np.float32).reshape(-1, 8)
    cam = np.ascontiguousarray(cam12, np.float32)
    li = np.ascontiguousarray(light3, np.float32)
    am = np.ascontiguousarray(ambient3, np.float32)
    img = np.empty((h, w, 3), np.float32)
    rays = C.c_uint64(0)
    lib = lib or load()
    _unpin()
    rc = lib.orc_rt_render(sph.ctypes.data, len(sph), pl.ctypes.data, len(pl), cam.ctypes.data, li.ctypes.data,
                              am.ctypes.data, W, H, x0, y0, w, h, max_depth, threads, img.ctypes.data, C.byref(rays))
    if rc != 0:
        raise ValueError("orc_rt_render: bad region")
    return img, rays.value
