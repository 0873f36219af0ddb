"""GPU parity of the queue-driven BFS (bfs_bench.cpp:327-633 semantics)
through the C ABI: every depth array must equal the serial BFS oracle's
bit for bit, and reached / edges scanned / levels must match it too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = [0, 1, 2, 3]
NAMES = {0: "gwfq", 1: "glfq", 2: "wfq", 3: "sfq"}
MODES = {0: "hostloop", 1: "persistent", 2: "async", 3: "bitmap"}
RING_KINDS = [0, 1, 3]  # the barrier-free mode takes ring frontiers


def _csr(nv, u, v):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    key = np.unique(np.concatenate([u * nv + v, v * nv + u]))
    s, t = np.divmod(key, nv)
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.int32), t.astype(np.int32)


def _run_and_check(orc, g, src, kind, threads, mode, off=None, idx=None):
    print(f"[bfs] kind={NAMES[kind]} mode={MODES[mode]} threads={threads} src={src} nv={g.info()[0]}", flush=True)
    if off is None:
        off, idx = g.csr()
    d, res = g.bfs(src, kind=kind, threads=threads, block=256 if threads >= 256 else 64, mode=mode)
    ref, reached, es = orc.bfs_serial(off, idx, src)
    bad = np.flatnonzero(d != ref)
    assert len(bad) == 0, f"{len(bad)} depths differ, first at {bad[:5]}: got {d[bad[:5]]} want {ref[bad[:5]]}"
    assert res["reached"] == reached
    if mode == 2:  # every expansion counts, a vertex expanded again after its depth fell included
        assert res["edges_scanned"] >= es
    else:
        assert res["edges_scanned"] == es
    assert res["levels"] == int(ref.max()) + 1
    return res


@pytest.mark.parametrize("mode", [0, 1, 2, 3], ids=[MODES[m] for m in (0, 1, 2, 3)])
@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_lattice_small(gcq, orc, kind, mode):
    if mode == 2 and kind == 2:
        pytest.skip("the barrier-free mode takes ring frontiers (WFQ is rejected: test_async_rejects_wfq)")
    from gcq_amd import bfs
    with bfs.Graph.lattice(64, 64, 0.75, 0.01, 7) as g:
        off, idx = g.csr()
        for threads in (64, 1024):
            _run_and_check(orc, g, 0, kind, threads, mode, off, idx)
        _run_and_check(orc, g, g.pick_source(5, 0), kind, 1024, mode, off, idx)


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_special_graphs(gcq, orc, kind):
    from gcq_amd import bfs
    n = 3000
    a = np.arange(n - 1)
    cases = {
        "path": (_csr(n, a, a + 1), 0),
        "path_mid": (_csr(n, a, a + 1), n // 2),
        "star": (_csr(n, np.zeros(n - 1, int), np.arange(1, n)), 7),  # hub of degree 2999 reached at depth 1
        "two_parts": (_csr(8, [0, 1, 4, 5], [1, 2, 5, 6]), 4),
        "isolated_src": (_csr(5, [0, 1], [1, 2]), 4),
        "single": ((np.array([0, 0], np.int32), np.array([], np.int32)), 0),
        "src_out_of_range": (_csr(6, [0, 1, 2], [1, 2, 3]), 1000),  # falls back to 0 (bfs_bench.cpp:555)
    }
    for name, ((off, idx), src) in cases.items():
        print(f"[bfs special] {name}", flush=True)
        with bfs.Graph.from_csr(off, idx) as g:
            for mode in ((0, 1, 2, 3) if kind in RING_KINDS else (0, 1, 3)):
                _run_and_check(orc, g, src, kind, 1024, mode, off, idx)


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_frontier_ring_of_exactly_v_slots(gcq, orc, kind):
    # V = 4,096 with 1,024 threads: the frontier rings hold exactly V slots.
    # The star pushes V - 1 vertices into one ring in one level, the path
    # fills both rings over V levels, the complete binary tree in 12 levels;
    # each graph runs twice, so the second run reuses the reset rings.  The
    # SFQ frontier's persistent and bitmap modes rely on every ticket staying
    # on the ring's first lap here.
    from gcq_amd import bfs
    n = 4096
    a = np.arange(n - 1)
    cases = {
        "star": (_csr(n, np.zeros(n - 1, int), np.arange(1, n)), 0),
        "path": (_csr(n, a, a + 1), 0),
        "tree": (_csr(n, (np.arange(1, n) - 1) // 2, np.arange(1, n)), 0),
    }
    for name, ((off, idx), src) in cases.items():
        print(f"[bfs V-slot rings] {name}", flush=True)
        with bfs.Graph.from_csr(off, idx) as g:
            for mode in (1, 3, 1):
                _run_and_check(orc, g, src, kind, 1024, mode, off, idx)


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_rmat_hubs(gcq, orc, kind):
    # Scale 15 R-MAT: frontiers well above the reference's 32,768-slot
    # rings and hub vertices of degree in the thousands.
    from gcq_amd import bfs
    with bfs.Graph.rmat(15, 16, 0.57, 0.19, 0.19, 11) as g:
        off, idx = g.csr()
        for k in range(2):
            _run_and_check(orc, g, g.pick_source(11, k), kind, 8192, 1, off, idx)
        _run_and_check(orc, g, g.pick_source(11, 2), kind, 8192, 0, off, idx)
        _run_and_check(orc, g, g.pick_source(11, 4), kind, 8192, 3, off, idx)
        if kind in RING_KINDS:
            _run_and_check(orc, g, g.pick_source(11, 3), kind, 8192, 2, off, idx)


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_lattice_1000_persistent_reuse(gcq, orc, kind):
    # One graph, repeated runs reuse the frontier queues (reset in place).
    from gcq_amd import bfs
    with bfs.Graph.lattice(1000, 1000, 0.75, 0.01, 7) as g:
        off, idx = g.csr()
        for src in (0, 999_999, 500_500):
            _run_and_check(orc, g, src, kind, 8192, 1, off, idx)


@pytest.mark.parametrize("mode", [1, 0], ids=[MODES[m] for m in (1, 0)])
def test_wfq_cells_recycled_65536_threads(gcq, orc, mode):
    # ~2,000 levels x 65,536 lanes would burn ~10^8 WFQ cells without the
    # three-frontier recycling; the pool holds V + 4 T + 2^20 cells.
    from gcq_amd import bfs
    with bfs.Graph.lattice(1000, 1000, 0.75, 0.01, 7) as g:
        _run_and_check(orc, g, 0, 2, 65536, mode)


@pytest.mark.parametrize("kind", [1, 0, 2, 3], ids=["glfq", "gwfq", "wfq", "sfq"])
def test_config3_lattice_4900(gcq, orc, kind):
    # The benchmark graph: 4900 x 4900 lattice, p = 0.75, 1% diagonals,
    # seed 7, source 0, 8,192 threads (block 256): persistent (9,371 levels
    # of counted pops, each level's Head set back at its barrier), the same
    # with visited-bitmap claims and, for the ring frontiers, barrier-free.
    from gcq_amd import bfs
    with bfs.Graph.lattice(4900, 4900, 0.75, 0.01, 7) as g:
        off, idx = g.csr()
        for mode in ((1, 3, 2) if kind in RING_KINDS else (1, 3)):
            res = _run_and_check(orc, g, 0, kind, 8192, mode, off, idx)
            print(f"[bfs cfg3] mode={MODES[mode]} levels={res['levels']} reached={res['reached']} "
                  f"kernel_ms={res['kernel_ms']:.1f} scanned={res['edges_scanned']}", flush=True)


def test_async_rejects_wfq(gcq):
    from gcq_amd import GcqError, bfs
    with bfs.Graph.lattice(16, 16, 0.75, 0.01, 7) as g:
        with pytest.raises(GcqError, match="ring frontier"):
            g.bfs(0, kind=2, threads=256, mode=2)


def test_kind_switch_same_graph(gcq, orc):
    from gcq_amd import bfs
    with bfs.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 11) as g:
        off, idx = g.csr()
        src = g.pick_source(3, 0)
        for kind in (0, 3, 1, 2, 0):
            _run_and_check(orc, g, src, kind, 2048, 1, off, idx)


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_graph_from_file(gcq, orc, kind, tmp_path):
    # A graph read from a SNAP edge list (bfs_bench.cpp:226-243 rules) runs
    # like any other CSR: lattice 120 x 150, one direction per line, 1-based.
    from gcq_amd import bfs
    off, idx = bfs.lattice(120, 150, keep=0.75, diag=0.01, seed=3)
    nv = len(off) - 1
    src = np.repeat(np.arange(nv), np.diff(off))
    up = src < idx
    path = tmp_path / "lattice.txt"
    path.write_text("# lattice\n" + "".join(f"{u + 1} {v + 1}\n" for u, v in zip(src[up], idx[up])))
    with bfs.Graph.load(str(path)) as g:
        o2, i2 = g.csr()
        if np.diff(off)[0] > 0 and np.diff(off)[-1] > 0:  # ids 1..nv all present: the same CSR
            assert np.array_equal(o2, off) and np.array_equal(i2, idx)
        _run_and_check(orc, g, 0, kind, 2048, 1, o2, i2)
