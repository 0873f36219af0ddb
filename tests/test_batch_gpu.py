"""GPU parity of the batched BFS: every query's depth array, reached count,
edges scanned and level count must equal the serial BFS oracle's, for every
queue kind, configuration and graph-replay state, with one work item per
(query, vertex) and with multi-source groups (up to 64 queries on one graph
as one bit-parallel traversal)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = {0: "gwfq", 1: "glfq", 2: "wfq", 3: "sfq"}


def _csr(nv, u, v):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    key = np.unique(np.concatenate([u * nv + v, v * nv + u]))
    s, t = np.divmod(key, nv)
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.int32), t.astype(np.int32)


def _mixed_graphs():
    from gcq_amd import bfs
    n = 2000
    a = np.arange(n - 1)
    return [
        bfs.Graph.lattice(64, 64, 0.75, 0.01, 7),
        bfs.Graph.rmat(10, 16, 0.57, 0.19, 0.19, 11),
        bfs.Graph.from_csr(*_csr(n, a, a + 1)),                           # path: 2,000 levels
        bfs.Graph.from_csr(*_csr(n, np.zeros(n - 1, int), np.arange(1, n))),  # star
        bfs.Graph.from_csr(*_csr(6, [0, 1], [1, 2])),                     # isolated vertices
    ]


def _check_all(orc, graphs, csrs, gids, srcs, outs, per):
    for i, (g, s) in enumerate(zip(gids, srcs)):
        off, idx = csrs[g]
        ref, reached, es = orc.bfs_serial(off, idx, int(s))
        d = outs[i]
        bad = np.flatnonzero(d != ref)
        assert len(bad) == 0, f"query {i} (graph {g}, src {s}): {len(bad)} depths differ, first {bad[:5]}"
        assert per[i]["reached"] == reached, f"query {i}"
        assert per[i]["edges_scanned"] == es, f"query {i}"
        assert per[i]["levels"] == int(ref.max()) + 1, f"query {i}"
        assert per[i]["status"] == 0


MS = [0, 1]
MS_IDS = ["per_query", "multi_source"]


@pytest.mark.parametrize("ms", MS, ids=MS_IDS)
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=[NAMES[k] for k in range(4)])
def test_batch_mixed_graphs(gcq, orc, kind, ms):
    from gcq_amd import bfs
    graphs = _mixed_graphs()
    csrs = [g.csr() for g in graphs]
    rng = np.random.default_rng(kind)
    gids = rng.integers(0, len(graphs), 40).astype(np.int32)
    srcs = np.array([rng.integers(0, graphs[g].info()[0]) for g in gids], np.int32)
    srcs[0] = 10**6  # out of range -> vertex 0 (bfs_bench.cpp:555)
    gids[1], srcs[1] = 4, 5  # isolated source
    print(f"[batch mixed] kind={NAMES[kind]} multi_source={ms}", flush=True)
    with bfs.Batch(graphs) as b:
        for use_graph in (0, 1, 1):
            outs, per, res = b.run(gids, srcs, depths=True, kind=kind, threads=4096, use_graph=use_graph,
                                   multi_source=ms)
            _check_all(orc, graphs, csrs, gids, srcs, outs, per)
            assert res["levels"] == max(p["levels"] for p in per)
    for g in graphs:
        g.close()


@pytest.mark.parametrize("knobs", [dict(pop_width=16), dict(pop_width=32, flush_at=16), dict(flush_at=8, block=128),
                                   dict(block=512, threads=8192), dict(capacity_log2=22), dict(threads=64, block=64)],
                         ids=["pw16", "pw32_fa16", "fa8_b128", "b512", "cap22", "one_wave"])
@pytest.mark.parametrize("ms", MS, ids=MS_IDS)
def test_batch_knobs(gcq, orc, knobs, ms):
    from gcq_amd import bfs
    graphs, gids, srcs = bfs.rmat_query_set(nq=48, scale_lo=8, scale_hi=13)
    csrs = [g.csr() for g in graphs]
    print(f"[batch knobs] {knobs} multi_source={ms}", flush=True)
    with bfs.Batch(graphs) as b:
        outs, per, res = b.run(gids, srcs, depths=True, multi_source=ms, **knobs)
        _check_all(orc, graphs, csrs, gids, srcs, outs, per)


@pytest.mark.parametrize("ms", MS, ids=MS_IDS)
def test_batch_graph_replay_new_sources(gcq, orc, ms):
    # A replayed graph must pick up new sources and graph choices (they live
    # in the slot and group tables uploaded before each launch).
    from gcq_amd import bfs
    graphs, gids, srcs = bfs.rmat_query_set(nq=32, scale_lo=9, scale_hi=12)
    csrs = [g.csr() for g in graphs]
    with bfs.Batch(graphs) as b:
        _, _, r0 = b.run(gids, srcs, multi_source=ms)
        assert r0["graph_replayed"] == 0
        rng = np.random.default_rng(5)
        for k in range(3):
            srcs2 = np.array([graphs[g].pick_source(100 + k, i) for i, g in enumerate(gids)], np.int32)
            outs, per, r = b.run(gids, srcs2, depths=True, multi_source=ms)
            assert r["graph_replayed"] == 1
            _check_all(orc, graphs, csrs, gids, srcs2, outs, per)
        gids3 = rng.permutation(gids).astype(np.int32)
        srcs3 = np.zeros_like(gids3)
        outs, per, r = b.run(gids3, srcs3, depths=True, multi_source=ms)
        _check_all(orc, graphs, csrs, gids3, srcs3, outs, per)


def test_batch_multi_source_groups_split(gcq, orc):
    # 150 queries on one graph make three groups (64 + 64 + 22); repeated
    # sources share a level-0 item; a second graph's queries interleave.
    from gcq_amd import bfs
    graphs = [bfs.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 11), bfs.Graph.lattice(40, 40, 0.75, 0.01, 7)]
    csrs = [g.csr() for g in graphs]
    rng = np.random.default_rng(3)
    gids = np.zeros(170, np.int32)
    gids[rng.choice(170, 20, replace=False)] = 1
    srcs = np.array([graphs[g].pick_source(9, i) for i, g in enumerate(gids)], np.int32)
    srcs[5] = srcs[6] = srcs[7]
    print("[batch ms groups] 170 queries, 150 on one graph", flush=True)
    with bfs.Batch(graphs) as b:
        for use_graph in (0, 1, 1):
            outs, per, res = b.run(gids, srcs, depths=True, multi_source=1, threads=8192, use_graph=use_graph)
            _check_all(orc, graphs, csrs, gids, srcs, outs, per)
    for g in graphs:
        g.close()


@pytest.mark.parametrize("ms", MS, ids=MS_IDS)
def test_batch_config4_full(gcq, orc, ms):
    # 256 queries, R-MAT scales 12..20, edge factor 16, seed 11.
    from gcq_amd import bfs
    graphs, gids, srcs = bfs.rmat_query_set()
    csrs = [g.csr() for g in graphs]
    with bfs.Batch(graphs) as b:
        outs, per, res = b.run(gids, srcs, depths=True, multi_source=ms, threads=65536)
        _check_all(orc, graphs, csrs, gids, srcs, outs, per)
        print(f"[batch cfg4] multi_source={ms} total_ms={res['total_ms']:.2f} kernel_ms={res['kernel_ms']:.2f} "
              f"levels={res['levels']}", flush=True)


def test_batch_autotune_small(gcq, orc):
    from gcq_amd import bfs
    graphs, gids, srcs = bfs.rmat_query_set(nq=32, scale_lo=8, scale_hi=11)
    csrs = [g.csr() for g in graphs]
    with bfs.Batch(graphs) as b:
        best, trials = b.autotune(gids, srcs, reps=1)
        assert len(trials) >= 5
        assert best.use_graph == 1
        outs, per, _ = b.run(gids, srcs, cfg=best, depths=True)
        _check_all(orc, graphs, csrs, gids, srcs, outs, per)
