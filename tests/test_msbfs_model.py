"""CPU model of the multi-source (bit-parallel) level protocol of the
batched BFS kernel (bfs_batch.hpp, ms_expand): groups of up to 64 queries on
one graph, a (group, vertex) work item per vertex and level, per-vertex
seen / next-level masks gathered by fetch-or.  The model processes each
level's items in a random order and lets every expansion read a neighbour's
seen word either as it stood at the level's start or as it stands now (the
kernel's cached reads may be stale), and checks every query's depths
against the serial BFS oracle.  A planted mistake (the gather word is not
cleared after its vertex settles) must be caught."""
import numpy as np
import pytest


def _csr(nv, u, v):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    key = np.unique(np.concatenate([u * nv + v, v * nv + u]))
    s, t = np.divmod(key, nv)
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.int32), t.astype(np.int32)


def ms_bfs_model(off, idx, srcs, rng, clear_gather=True):
    """Depth arrays of the queries srcs (one group of <= 64), -1 unreached."""
    nv = len(off) - 1
    m = len(srcs)
    assert m <= 64
    seen = [0] * nv
    nxt = [[0] * nv, [0] * nv]
    depth = np.full((m, nv), -1, np.int64)
    frontier = []
    for b, s in enumerate(srcs):  # seeding: level 0 reads nxt[1]
        if nxt[1][s] == 0:
            frontier.append(s)
        nxt[1][s] |= 1 << b
    L = 0
    while frontier:
        assert len(frontier) == len(set(frontier)), "a vertex queued twice in one level"
        start_seen = list(seen)
        out = []
        rng.shuffle(frontier)
        for v in frontier:
            arrived = nxt[(L + 1) & 1][v]
            if clear_gather:
                nxt[(L + 1) & 1][v] = 0
            fresh = arrived & ~seen[v]
            if not fresh:
                continue
            seen[v] |= fresh
            for b in range(m):
                if fresh >> b & 1:
                    depth[b, v] = L
            for e in range(off[v], off[v + 1]):
                w = idx[e]
                sw = start_seen[w] if rng.random() < 0.5 else seen[w]  # a cached read may be stale
                send = fresh & ~sw
                if send:
                    old = nxt[L & 1][w]
                    nxt[L & 1][w] |= send
                    if old == 0:
                        out.append(w)
        frontier = out
        L += 1
        assert L <= nv + 1, "no progress"
    return depth


def _graphs(rng):
    n = 400
    u = rng.integers(0, n, 1500)
    v = rng.integers(0, n, 1500)
    yield "random", _csr(n, u, v)
    a = np.arange(299)
    yield "path", _csr(300, a, a + 1)
    side = 20
    r, c = np.divmod(np.arange(side * side), side)
    right = np.flatnonzero(c < side - 1)
    down = np.flatnonzero(r < side - 1)
    keep = rng.random(len(right) + len(down)) < 0.75
    uu = np.concatenate([right, down])[keep]
    vv = np.concatenate([right + 1, down + side])[keep]
    yield "lattice", _csr(side * side, uu, vv)
    yield "two_parts", _csr(8, [0, 1, 4, 5], [1, 2, 5, 6])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_model_matches_serial_bfs(orc, seed):
    rng = np.random.default_rng(seed)
    for name, (off, idx) in _graphs(rng):
        nv = len(off) - 1
        srcs = [int(x) for x in rng.integers(0, nv, 64)]
        srcs[1] = srcs[0]  # two members with the same source
        d = ms_bfs_model(off, idx, srcs, rng)
        for b, s in enumerate(srcs):
            ref, _, _ = orc.bfs_serial(off, idx, s)
            assert np.array_equal(d[b], ref), f"{name}: member {b} (source {s}) differs"


def test_planted_missing_clear_is_caught(orc):
    # Without clearing a settled vertex's gather word, bits of level L linger
    # into level L+2, and the vertex is never queued again by a fetch-or that
    # finds the word non-zero: depths go wrong (or the frontier repeats).
    rng = np.random.default_rng(7)
    caught = False
    for name, (off, idx) in _graphs(rng):
        nv = len(off) - 1
        srcs = [int(x) for x in rng.integers(0, nv, 16)]
        try:
            d = ms_bfs_model(off, idx, srcs, rng, clear_gather=False)
        except AssertionError:
            caught = True
            continue
        for b, s in enumerate(srcs):
            ref, _, _ = orc.bfs_serial(off, idx, s)
            if not np.array_equal(d[b], ref):
                caught = True
    assert caught
