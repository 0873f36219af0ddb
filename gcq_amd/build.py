"""Build libgcq.so (HIP, gfx950) in-tree.

The library is compiled by hipcc directly from gcq_amd/csrc/*.hip; the
resulting gcq_amd/libgcq.so is what every Python entry point loads.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgcq.so")
ARCH = os.environ.get("GCQ_OFFLOAD_ARCH", "gfx950")

# Every HIP TU of the library.
SOURCES = ["core.hip", "bfs.hip", "batch.hip", "graph.cpp", "rt.hip", "rt_compact.hip", "serve.hip", "dropin_check.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-pthread",
    f"-I{CSRC}",  # the drop-in headers in include/ reach the queue code as "gcq/..."
    # packed-FP32 forms are disabled for all device code (known-bad forms on
    # these boxes); hipcc's host pass warns that it does not know the feature.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, jobs=None):
    srcs = sources()
    objs = [os.path.join(CSRC, "_obj", os.path.splitext(s)[0] + ".o") for s in srcs]
    os.makedirs(os.path.join(CSRC, "_obj"), exist_ok=True)
    cmds = [["hipcc", *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(srcs, objs)]
    # one hipcc per translation unit, at most one per core (capped at 8)
    jobs = jobs or max(1, min(8, os.cpu_count() or 1, len(cmds)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    if verbose and out.strip():
        print(out)


if __name__ == "__main__":
    print(build(verbose=True))
