"""ORACLE (test infrastructure only): a plain-Python restatement of the
reference's graph-file loaders, used to check the library's loader on
generated files.

load_mtx  follows concurrent_bfs/bfs_bench.cpp:202-224
load_snap follows concurrent_bfs/bfs_bench.cpp:226-243
csr       follows concurrent_bfs/bfs_bench.cpp:185-200 (csr_from_edges)
Where the reference would return an empty graph (no banner, no size line,
no vertices) these raise ValueError, as the library reports an error there.
"""
import re

import numpy as np

_INT = re.compile(r"[ \t\r\v\f]*([+-]?[0-9]+)")


def _two_ints(line, k=2):
    """The first k integers of a line, parsed as `istream >> long long`
    does (leading blanks, optional sign, digits; stop at anything else)."""
    out, pos = [], 0
    for _ in range(k):
        m = _INT.match(line, pos)
        if not m:
            return None
        v = int(m.group(1))
        if not -(1 << 63) <= v < (1 << 63):
            return None
        out.append(v)
        pos = m.end()
    return out


def csr(nv, edges):
    """Sorted, deduplicated CSR; pairs with an endpoint outside [0, nv) dropped."""
    keep = sorted({(u, v) for u, v in edges if 0 <= u < nv and 0 <= v < nv})
    off = np.zeros(nv + 1, dtype=np.int64)
    for u, _ in keep:
        off[u + 1] += 1
    off = np.cumsum(off).astype(np.int32)
    idx = np.array([v for _, v in keep], dtype=np.int32)
    return off, idx


def load_mtx(text, undirected=True):
    lines = text.split("\n")
    if not lines or not lines[0].startswith("%%MatrixMarket"):
        raise ValueError("no banner")
    size, k = None, 1
    while k < len(lines):
        ln = lines[k]
        k += 1
        if ln == "" or ln[0] == "%":
            continue
        size = _two_ints(ln, 3)
        break
    if size is None or min(size) < 0:
        raise ValueError("bad size line")
    nr, nc, nz = size
    edges, cnt = [], 0
    while cnt < nz and k < len(lines):
        ln = lines[k]
        k += 1
        if ln == "" or ln[0] == "%":
            continue
        ij = _two_ints(ln)
        if ij is None:
            continue
        i, j = ij[0] - 1, ij[1] - 1
        cnt += 1
        if i == j or i < 0 or j < 0 or i >= nr or j >= nc:
            continue
        edges.append((i, j))
        if undirected:
            edges.append((j, i))
    nv = max(nr, nc)
    if nv <= 0:
        raise ValueError("no vertices")
    return csr(nv, edges)


def load_snap(text, undirected=True):
    edges, mn, mx = [], 1 << 60, -1
    for ln in text.split("\n"):
        if ln == "" or ln[0] in "#%":
            continue
        uv = _two_ints(ln)
        if uv is None or uv[0] == uv[1]:
            continue
        u, v = uv
        mn, mx = min(mn, u, v), max(mx, u, v)
        edges.append((u, v))
        if undirected:
            edges.append((v, u))
    if mn == 1:
        edges = [(u - 1, v - 1) for u, v in edges]
        mx -= 1
    nv = mx + 1
    if nv <= 0:
        raise ValueError("no vertices")
    return csr(nv, edges)
