"""ctypes bindings for the workload entry points of libgcq (graphs, BFS,
batched/served BFS, ray tracing).  Registered by gcq_amd.load()."""
import ctypes as C


class BfsConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("threads", C.c_uint32), ("block", C.c_uint32), ("mode", C.c_int32)]


class BfsResult(C.Structure):
    _fields_ = [("edges_scanned", C.c_uint64), ("reached", C.c_uint64), ("levels", C.c_int32), ("pad", C.c_int32),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double)]


class BatchConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("threads", C.c_uint32), ("block", C.c_uint32), ("pop_width", C.c_uint32),
                ("flush_at", C.c_uint32), ("capacity_log2", C.c_uint32), ("use_graph", C.c_uint32),
                ("multi_source", C.c_uint32)]


class QueryResult(C.Structure):
    _fields_ = [("reached", C.c_uint64), ("edges_scanned", C.c_uint64), ("levels", C.c_int32),
                ("status", C.c_int32)]


class BatchResult(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("kernel_ms", C.c_double), ("edges_scanned", C.c_uint64),
                ("levels", C.c_int32), ("graph_replayed", C.c_int32)]


class TuneTrial(C.Structure):
    _fields_ = [("cfg", BatchConfig), ("ms", C.c_double)]


class RtSceneDesc(C.Structure):
    _fields_ = [("spheres", C.c_void_p), ("nspheres", C.c_int32), ("nplanes", C.c_int32), ("planes", C.c_void_p),
                ("cam_o", C.c_float * 3), ("cam_f", C.c_float * 3), ("cam_r", C.c_float * 3), ("cam_u", C.c_float * 3),
                ("light", C.c_float * 3), ("ambient", C.c_float * 3)]


class RtConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("threads_per_tile", C.c_uint32), ("block", C.c_uint32), ("max_depth", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
                ("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32)]


class RtResult(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("total_rays", C.c_uint64), ("kernel_ms", C.c_double),
                ("total_ms", C.c_double), ("tiles", C.c_int32), ("pad", C.c_int32)]


class RtCsvRow(C.Structure):
    _fields_ = [("gpu", C.c_char_p), ("queue", C.c_char_p), ("scene", C.c_char_p), ("threads", C.c_uint32),
                ("bounces", C.c_int32), ("res", C.POINTER(RtResult))]


class Request(C.Structure):
    _fields_ = [("type", C.c_int32), ("graph", C.c_int32), ("src", C.c_int32), ("x0", C.c_int32), ("y0", C.c_int32),
                ("w", C.c_int32), ("h", C.c_int32), ("keep", C.c_int32), ("id", C.c_uint64), ("arrival_ms", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("type", C.c_int32), ("status", C.c_int32), ("levels", C.c_int32),
                ("pad", C.c_int32), ("rays", C.c_uint64), ("arrival_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("bfs_threads", C.c_uint32), ("rt_threads", C.c_uint32), ("block", C.c_uint32),
                ("bfs_slots", C.c_uint32), ("rt_slots", C.c_uint32), ("frontier_log2", C.c_uint32),
                ("max_depth", C.c_int32), ("pop_width", C.c_uint32), ("flush_at", C.c_uint32),
                ("budget_bytes", C.c_uint64), ("keep_bytes", C.c_uint64), ("max_requests", C.c_uint32),
                ("small_log2", C.c_uint32), ("small_threads", C.c_uint32), ("multi_source", C.c_uint32)]


_P = C.POINTER
VP = C.c_void_p

SIGNATURES = {
    "gcq_graph_lattice": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64, _P(VP)]),
    "gcq_graph_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, _P(VP)]),
    "gcq_graph_from_csr": (C.c_int, [C.c_int32, C.c_int64, VP, VP, _P(VP)]),
    "gcq_graph_load": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _P(VP)]),
    "gcq_graph_info": (C.c_int, [VP, _P(C.c_int32), _P(C.c_int64)]),
    "gcq_graph_copy_csr": (C.c_int, [VP, VP, VP]),
    "gcq_graph_pick_source": (C.c_int32, [VP, C.c_uint64, C.c_uint64]),
    "gcq_graph_destroy": (C.c_int, [VP]),
    "gcq_bfs_run": (C.c_int, [VP, C.c_int32, _P(BfsConfig), VP, _P(BfsResult)]),
    "gcq_bfs_run_noalloc": (C.c_int, [VP, VP, C.c_int32, C.c_int64, C.c_int32, C.c_uint32, C.c_uint32, VP, VP,
                                      _P(BfsResult)]),
    "gcq_batch_create": (C.c_int, [_P(VP), C.c_int32, _P(VP)]),
    "gcq_batch_config_default": (None, [_P(BatchConfig)]),
    "gcq_batch_run": (C.c_int, [VP, _P(BatchConfig), C.c_int32, VP, VP, _P(VP), _P(QueryResult), _P(BatchResult)]),
    "gcq_batch_autotune": (C.c_int, [VP, C.c_int32, C.c_int32, VP, VP, C.c_int32, _P(BatchConfig), _P(TuneTrial),
                                     C.c_int32, _P(C.c_int32)]),
    "gcq_batch_destroy": (C.c_int, [VP]),
    "gcq_rt_scene_create": (C.c_int, [_P(RtSceneDesc), _P(VP)]),
    "gcq_rt_scene_preset": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, _P(VP)]),
    "gcq_rt_scene_get": (C.c_int, [VP, _P(RtSceneDesc)]),
    "gcq_rt_scene_destroy": (C.c_int, [VP]),
    "gcq_rt_config_default": (None, [_P(RtConfig)]),
    "gcq_rt_render": (C.c_int, [VP, _P(RtConfig), VP, _P(RtResult)]),
    "gcq_rt_render_compact": (C.c_int, [VP, _P(RtConfig), VP, _P(RtResult)]),
    "gcq_rt_write_ppm": (C.c_int, [VP, C.c_int32, C.c_int32, C.c_char_p]),
    "gcq_rt_append_csv": (C.c_int, [C.c_char_p, C.c_int32, _P(RtCsvRow)]),
    "gcq_server_config_default": (None, [_P(ServerConfig)]),
    "gcq_server_create": (C.c_int, [_P(VP), C.c_int32, VP, C.c_int32, C.c_int32, _P(ServerConfig), _P(VP)]),
    "gcq_server_start": (C.c_int, [VP]),
    "gcq_server_submit": (C.c_int, [VP, _P(Request)]),
    "gcq_server_poll": (C.c_int, [VP, _P(Completion), C.c_int32, _P(C.c_int32)]),
    "gcq_server_stop": (C.c_int, [VP]),
    "gcq_server_fetch": (C.c_int, [VP, C.c_uint64, VP]),
    "gcq_server_run_trace": (C.c_int, [VP, _P(Request), C.c_int32, _P(Completion)]),
    "gcq_server_destroy": (C.c_int, [VP]),
    "gcq_plain_run_trace": (C.c_int, [_P(VP), C.c_int32, VP, C.c_int32, C.c_int32, _P(BfsConfig), _P(RtConfig),
                                      _P(Request), C.c_int32, _P(Completion)]),
}


def register(lib):
    """Attach argument/result types for the workload symbols present in lib."""
    for name, (res, args) in SIGNATURES.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib
