"""ctypes bindings for the workload entry points of libgcq (BFS, ray
tracing, serving).  Registered by gcq_amd.load()."""


def register(lib):
    """Attach argument/result types for the workload symbols present in lib."""
    return lib
