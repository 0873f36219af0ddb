"""gcq_amd — MI355X-native GPU MPMC FIFO queues (G-WFQ, G-LFQ, WFQ, SFQ) and
the queue-driven BFS / ray-compaction workloads, behind the C ABI of
libgcq.so (include/gcq.h).

This module is host-side plumbing only: it loads the in-tree libgcq.so with
ctypes and mirrors the C structs.  There is no CPU fallback; if the library
is missing every call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcq.so")

GWFQ, GLFQ, WFQ, SFQ = 0, 1, 2, 3
KIND_NAMES = {GWFQ: "G-WFQ", GLFQ: "G-LFQ", WFQ: "WFQ", SFQ: "SFQ"}
KINDS = {"gwfq": GWFQ, "glfq": GLFQ, "wfq": WFQ, "sfq": SFQ}
BALANCED, SPLIT = 0, 1


class GcqError(RuntimeError):
    def __init__(self, rc, where):
        lib = _lib_or_none()
        msg = lib.gcq_last_error().decode() if lib else ""
        name = lib.gcq_strerror(rc).decode() if lib else str(rc)
        super().__init__(f"{where}: {name} ({rc}) {msg}")
        self.rc = rc


class QueueConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("capacity", C.c_uint32), ("max_threads", C.c_uint32),
                ("remap", C.c_uint32), ("help_delay", C.c_uint32), ("enq_patience", C.c_uint32),
                ("deq_patience", C.c_uint32), ("wfq_segment_cells", C.c_uint32), ("poll_backoff", C.c_uint32)]


class FaaConfig(C.Structure):
    _fields_ = [("threads", C.c_uint32), ("block", C.c_uint32), ("counters", C.c_uint32),
                ("stride_bytes", C.c_uint32), ("aggregate", C.c_uint32), ("iters", C.c_uint32),
                ("launches", C.c_uint32), ("poll_percent", C.c_uint32), ("poll_other_line", C.c_uint32)]


class FaaResult(C.Structure):
    _fields_ = [("tickets_per_s", C.c_double), ("atomics_per_s", C.c_double), ("kernel_ms", C.c_double),
                ("loads_per_s", C.c_double)]


class TbConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("threads", C.c_uint32), ("block", C.c_uint32), ("chunk_ops", C.c_uint32),
                ("producer_percent", C.c_uint32), ("warmup_ms", C.c_uint32), ("run_ms", C.c_uint32),
                ("launches_per_sync", C.c_uint32), ("warmup_launches", C.c_uint32), ("timed_launches", C.c_uint32),
                ("solo_claims", C.c_uint32)]


class TbResult(C.Structure):
    _fields_ = [("enq_success", C.c_uint64), ("enq_fail", C.c_uint64), ("deq_success", C.c_uint64),
                ("deq_empty", C.c_uint64), ("launches", C.c_uint64), ("elapsed_ms", C.c_double),
                ("kernel_ms_mean", C.c_double), ("succ_mops", C.c_double), ("pre_enq_success", C.c_uint64),
                ("pre_deq_success", C.c_uint64)]


class FifoResult(C.Structure):
    _fields_ = [("in_count", C.c_uint64), ("out_count", C.c_uint64), ("err_oob", C.c_uint32),
                ("err_dup", C.c_uint32), ("err_order", C.c_uint32), ("err_missing", C.c_uint32),
                ("err_gt1", C.c_uint32), ("verified", C.c_int32)]


class HistConfig(C.Structure):
    _fields_ = [("threads", C.c_uint32), ("ops", C.c_uint32), ("block", C.c_uint32), ("mode", C.c_uint32),
                ("producer_percent", C.c_uint32), ("delay_every", C.c_uint32), ("delay_cycles", C.c_uint32),
                ("seed", C.c_uint64)]


HIST_DTYPE = np.dtype([("proc", np.uint32), ("op", np.uint32), ("arg", np.uint64), ("ret", np.uint64),
                       ("call", np.uint64), ("end", np.uint64), ("steps", np.uint32), ("slow", np.uint32)])

_lib = None


def _lib_or_none():
    return _lib


def _P(t):
    return C.POINTER(t)


def load():
    """Load libgcq.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "gcq_strerror": (C.c_char_p, [C.c_int]),
        "gcq_last_error": (C.c_char_p, []),
        "gcq_version": (C.c_char_p, []),
        "gcq_device_info": (C.c_int, [C.c_char_p, C.c_size_t, _P(C.c_int), _P(C.c_uint64)]),
        "gcq_queue_config_default": (None, [C.c_int, _P(QueueConfig)]),
        "gcq_queue_create": (C.c_int, [_P(QueueConfig), _P(C.c_void_p)]),
        "gcq_queue_reset": (C.c_int, [C.c_void_p]),
        "gcq_queue_destroy": (C.c_int, [C.c_void_p]),
        "gcq_queue_device_ptr": (C.c_void_p, [C.c_void_p]),
        "gcq_queue_bytes": (C.c_uint64, [C.c_void_p]),
        "gcq_faa_peak": (C.c_int, [_P(FaaConfig), _P(FaaResult)]),
        "gcq_fence_cost": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P(C.c_double)]),
        "gcq_hbm_copy": (C.c_int, [C.c_uint64, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
        "gcq_round_trip": (C.c_int, [C.c_uint32, C.c_uint32, _P(C.c_double)]),
        "gcq_dropin_selfcheck": (C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, _P(C.c_uint32)]),
        "gcq_tb_run": (C.c_int, [C.c_void_p, _P(TbConfig), _P(TbResult)]),
        "gcq_queue_drain": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint64)]),
        "gcq_fifo_check": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _P(FifoResult)]),
        "gcq_hist_run": (C.c_int, [C.c_void_p, _P(HistConfig), C.c_void_p, C.c_size_t, _P(C.c_size_t)]),
        "gcq_hist_write_jsonl": (C.c_int, [C.c_void_p, C.c_size_t, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    _register_workload_symbols(lib)
    return lib


def _register_workload_symbols(lib):
    from . import workloads  # noqa: WPS433  (BFS / RT / serving bindings)
    workloads.register(lib)


def _check(rc, where):
    if rc != 0:
        raise GcqError(rc, where)


def version():
    return load().gcq_version().decode()


def device_info():
    lib = load()
    name = C.create_string_buffer(256)
    cu = C.c_int(0)
    hbm = C.c_uint64(0)
    _check(lib.gcq_device_info(name, 256, C.byref(cu), C.byref(hbm)), "gcq_device_info")
    return {"name": name.value.decode(), "cus": cu.value, "hbm_bytes": hbm.value}


def default_config(kind):
    cfg = QueueConfig()
    load().gcq_queue_config_default(kind, C.byref(cfg))
    return cfg


class Queue:
    """A device-resident MPMC queue (one of GWFQ, GLFQ, WFQ, SFQ)."""

    def __init__(self, kind, capacity=1 << 20, max_threads=65536, **kw):
        lib = load()
        cfg = default_config(kind)
        cfg.capacity = capacity
        cfg.max_threads = max_threads
        for k, v in kw.items():
            setattr(cfg, k, v)
        h = C.c_void_p()
        _check(lib.gcq_queue_create(C.byref(cfg), C.byref(h)), "gcq_queue_create")
        self.kind = kind
        self.cfg = cfg
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("queue destroyed")
        return self._h

    def reset(self):
        _check(load().gcq_queue_reset(self.handle), "gcq_queue_reset")

    def nbytes(self):
        return int(load().gcq_queue_bytes(self.handle))

    def device_ptr(self):
        return int(load().gcq_queue_device_ptr(self.handle) or 0)

    def close(self):
        if self._h is not None:
            load().gcq_queue_destroy(self._h)
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

    # ------------------------------------------------------------------ harnesses
    def tb_run(self, mode, threads, block=256, chunk_ops=128, producer_percent=25, warmup_ms=200, run_ms=1000,
               launches_per_sync=0, warmup_launches=0, timed_launches=0, solo_claims=False):
        """Timed balanced/split run: a time window (run_ms) or, with
        timed_launches > 0, exactly that many launches after warmup_launches.
        solo_claims turns off workgroup combining (G-WFQ, G-LFQ): of idle
        consumers' claims and of producers' Tail tickets."""
        cfg = TbConfig(mode, threads, block, chunk_ops, producer_percent, warmup_ms, run_ms, launches_per_sync,
                       warmup_launches, timed_launches, int(bool(solo_claims)))
        res = TbResult()
        _check(load().gcq_tb_run(self.handle, C.byref(cfg), C.byref(res)), "gcq_tb_run")
        return {f: getattr(res, f) for f, _ in TbResult._fields_}

    def drain(self, threads=4096):
        """Dequeue everything the (quiescent) queue holds; returns the count."""
        n = C.c_uint64(0)
        _check(load().gcq_queue_drain(self.handle, threads, C.byref(n)), "gcq_queue_drain")
        return n.value

    def fifo_check(self, threads, ops, block=256, concurrent=False):
        res = FifoResult()
        _check(load().gcq_fifo_check(self.handle, threads, ops, block, int(concurrent), C.byref(res)),
               "gcq_fifo_check")
        return {f: getattr(res, f) for f, _ in FifoResult._fields_}

    def hist_run(self, threads, ops, block=64, mode=0, producer_percent=50, delay_every=0, delay_cycles=0, seed=1):
        cfg = HistConfig(threads, ops, block, mode, producer_percent, delay_every, delay_cycles, seed)
        cap = threads * ops * (2 if mode == 0 else 1)
        out = np.zeros(cap, dtype=HIST_DTYPE)
        n = C.c_size_t(0)
        _check(load().gcq_hist_run(self.handle, C.byref(cfg), out.ctypes.data, cap, C.byref(n)), "gcq_hist_run")
        return out[: n.value]


def write_jsonl(entries, path):
    e = np.ascontiguousarray(entries, dtype=HIST_DTYPE)
    _check(load().gcq_hist_write_jsonl(e.ctypes.data, len(e), path.encode()), "gcq_hist_write_jsonl")


def faa_peak(threads=65536, block=256, counters=1, stride_bytes=128, aggregate=True, iters=256, launches=10,
             poll_percent=0, poll_other_line=False):
    """Counter-increment throughput; with poll_percent > 0 that share of the
    waves only loads (the counter, or another line) as idle pollers would."""
    cfg = FaaConfig(threads, block, counters, stride_bytes, int(aggregate), iters, launches, poll_percent,
                    int(poll_other_line))
    res = FaaResult()
    _check(load().gcq_faa_peak(C.byref(cfg), C.byref(res)), "gcq_faa_peak")
    return {f: getattr(res, f) for f, _ in FaaResult._fields_}


def fence_cost(mode, threads=65536, block=256, iters=512):
    ns = C.c_double(0)
    _check(load().gcq_fence_cost(threads, block, mode, iters, C.byref(ns)), "gcq_fence_cost")
    return ns.value


ROUND_TRIP_MODES = ("load", "atomic", "cached_load")


def round_trip(mode, steps=20000):
    """Nanoseconds per dependent device-scope access (mode: index or name in
    ROUND_TRIP_MODES) along a random cycle of 65,536 lines."""
    if isinstance(mode, str):
        mode = ROUND_TRIP_MODES.index(mode)
    ns = C.c_double(0)
    _check(load().gcq_round_trip(mode, steps, C.byref(ns)), "gcq_round_trip")
    return ns.value


def hbm_copy(nbytes=1 << 30, iters=20):
    """Measured streaming-copy bandwidth (GB/s, read + write) and launch time."""
    gbps, ms = C.c_double(0), C.c_double(0)
    _check(load().gcq_hbm_copy(nbytes, iters, C.byref(gbps), C.byref(ms)), "gcq_hbm_copy")
    return gbps.value, ms.value


DROPIN_APIS = ("gwf_ptr", "gwf_ring", "glf_ptr", "glf_ring", "wfq", "sfq_blocking", "sfq_nb")
DROPIN_ERRORS = ("order", "capacity", "not_empty", "duplicate", "missing", "spin_limit")


def dropin_selfcheck(api, threads=1024, ops=16):
    """Run the drop-in header API `api` (index or name in DROPIN_APIS); returns
    the error counters by name (all zero when the API behaves)."""
    if isinstance(api, str):
        api = DROPIN_APIS.index(api)
    errs = (C.c_uint32 * 6)()
    _check(load().gcq_dropin_selfcheck(api, threads, ops, errs), "gcq_dropin_selfcheck")
    return dict(zip(DROPIN_ERRORS, list(errs)))
