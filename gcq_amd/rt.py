"""Queue-compacted ray tracing (the reference's concurrent_rt workload,
rt_final.cpp) through libgcq's C ABI."""
import ctypes as C

import numpy as np

from . import GLFQ, _check, load  # noqa: F401
from .workloads import RtConfig, RtCsvRow, RtResult, RtSceneDesc

COMPLEX, CORNELL, RANDOM = 0, 1, 2
CSV_QUEUE, CSV_COMPACT = 0, 1
QUEUE_CSV_NAMES = {0: "GWFQ", 1: "GLFQ", 2: "WFQ", 3: "SFQ"}


def write_ppm(img, path):
    """8-bit binary PPM of an [h, w, 3] float image, truncated as the
    reference's save_ppm does (rt_final.cpp:689-700)."""
    a = np.ascontiguousarray(img, np.float32)
    h, w, _ = a.shape
    _check(load().gcq_rt_write_ppm(a.ctypes.data, w, h, str(path).encode()), "gcq_rt_write_ppm")


def append_csv(path, style, result, gpu, queue, scene="", threads=0, bounces=0):
    """Append one row (result: the dict render()/render_compact() return) in
    the reference's rt_queue_perf.csv (CSV_QUEUE) or rt_performance.csv
    (CSV_COMPACT) columns."""
    res = RtResult()
    for f, _ in RtResult._fields_:
        if f in result:
            setattr(res, f, result[f])
    row = RtCsvRow(gpu.encode(), queue.encode(), scene.encode(), threads, bounces, C.pointer(res))
    _check(load().gcq_rt_append_csv(str(path).encode(), style, C.byref(row)), "gcq_rt_append_csv")


def config(**kw):
    c = RtConfig()
    load().gcq_rt_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Scene:
    """Spheres and planes (8 floats each: centre|normal, radius|d, albedo,
    reflectivity), camera (origin, forward, right, up), light, ambient."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def preset(cls, which, nspheres=64, seed=3):
        h = C.c_void_p()
        _check(load().gcq_rt_scene_preset(which, nspheres, seed, C.byref(h)), "gcq_rt_scene_preset")
        return cls(h)

    @classmethod
    def create(cls, spheres, planes, cam_o, cam_f, cam_r=(1, 0, 0), cam_u=(0, 1, 0), light=(-0.5, 1, 0.2),
               ambient=(0.05, 0.05, 0.05)):
        sph = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 8))
        pl = np.ascontiguousarray(np.asarray(planes, np.float32).reshape(-1, 8))
        d = RtSceneDesc()
        d.spheres, d.nspheres = (sph.ctypes.data if len(sph) else None), len(sph)
        d.planes, d.nplanes = (pl.ctypes.data if len(pl) else None), len(pl)
        for name, v in (("cam_o", cam_o), ("cam_f", cam_f), ("cam_r", cam_r), ("cam_u", cam_u), ("light", light),
                        ("ambient", ambient)):
            getattr(d, name)[:] = [float(x) for x in v]
        h = C.c_void_p()
        _check(load().gcq_rt_scene_create(C.byref(d), C.byref(h)), "gcq_rt_scene_create")
        return cls(h)

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("scene destroyed")
        return self._h

    def arrays(self):
        """(spheres[n,8], planes[m,8], cam12, light3, ambient3) as float32."""
        d = RtSceneDesc()
        _check(load().gcq_rt_scene_get(self.handle, C.byref(d)), "gcq_rt_scene_get")
        sph = np.ctypeslib.as_array((C.c_float * (8 * d.nspheres)).from_address(d.spheres)).reshape(-1, 8).copy() \
            if d.nspheres else np.zeros((0, 8), np.float32)
        pl = np.ctypeslib.as_array((C.c_float * (8 * d.nplanes)).from_address(d.planes)).reshape(-1, 8).copy() \
            if d.nplanes else np.zeros((0, 8), np.float32)
        cam = np.array(list(d.cam_o) + list(d.cam_f) + list(d.cam_r) + list(d.cam_u), np.float32)
        return sph, pl, cam, np.array(list(d.light), np.float32), np.array(list(d.ambient), np.float32)

    def render(self, cfg=None, **kw):
        """Returns (image[h, w, 3] float32, result dict)."""
        cfg = cfg if cfg is not None else config(**kw)
        w = cfg.w or cfg.width
        h = cfg.h or cfg.height
        img = np.empty((h, w, 3), np.float32)
        res = RtResult()
        _check(load().gcq_rt_render(self.handle, C.byref(cfg), img.ctypes.data, C.byref(res)), "gcq_rt_render")
        return img, {f: getattr(res, f) for f, _ in RtResult._fields_ if f != "pad"}

    def render_compact(self, cfg=None, **kw):
        """The stream-compaction comparator (one launch per bounce, live rays
        packed by a prefix sum); same image as render().  Returns
        (image[h, w, 3] float32, result dict)."""
        cfg = cfg if cfg is not None else config(**kw)
        w = cfg.w or cfg.width
        h = cfg.h or cfg.height
        img = np.empty((h, w, 3), np.float32)
        res = RtResult()
        _check(load().gcq_rt_render_compact(self.handle, C.byref(cfg), img.ctypes.data, C.byref(res)),
               "gcq_rt_render_compact")
        return img, {f: getattr(res, f) for f, _ in RtResult._fields_ if f != "pad"}

    def close(self):
        if self._h is not None:
            load().gcq_rt_scene_destroy(self._h)
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
