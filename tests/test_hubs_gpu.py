"""GPU parity of long adjacencies in the persistent BFS (one traversal, in
both claim modes) and in the batched and served BFS (per-query items and
multi-source groups): a vertex with more than 4,096 edges is posted by the wave that settles it and
scanned by the whole grid in 1,024-edge chunks between two barriers of its
level.  The graph below makes every part of that path run:

- a centre joined to 2,100 hubs, so level 1 posts more vertices than the
  level's table holds (2,048): the rest are scanned by their own waves;
- a hub of degree exactly 4,096 (never posted) and one of 4,097 (posted;
  its last chunk holds one edge);
- a hub of 200,000 edges (196 chunks) and hubs of 4,097..4,399 edges;
- two hubs that share 3,000 leaves (one leaf reached from two chunks at the
  same level);
- a hub two levels further out (10,000 edges, reached through a leaf) that
  also reaches 2,000 leaves of another hub across levels.

Every depth array, reached count, edge count and level count must equal the
serial BFS oracle's."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def _hub_graph():
    rng = np.random.default_rng(5)
    H = 2100
    leaves = rng.integers(4097, 4400, H)
    leaves[0] = 4095      # hub 1: degree 4,096 with the centre edge
    leaves[1] = 4096      # hub 2: degree 4,097
    leaves[2] = 200_000   # hub 3
    first_leaf = 1 + H + np.concatenate([[0], np.cumsum(leaves)[:-1]])
    nv0 = 1 + H + int(leaves.sum())
    us, vs = [np.zeros(H, np.int64)], [np.arange(1, H + 1, dtype=np.int64)]
    hub_of_leaf = np.repeat(np.arange(1, H + 1), leaves)
    us.append(hub_of_leaf)
    vs.append(np.arange(1 + H, nv0, dtype=np.int64))
    # hubs 5 and 6 share 3,000 of hub 5's leaves
    us.append(np.full(3000, 6))
    vs.append(first_leaf[4] + np.arange(3000))
    # a second-level hub behind the first leaf of hub 4, with 10,000 own
    # leaves and 2,000 of hub 7's
    S = nv0
    nv = S + 1 + 10_000
    us += [np.array([first_leaf[3]]), np.full(10_000, S), np.full(2000, S)]
    vs += [np.array([S]), np.arange(S + 1, nv), first_leaf[6] + np.arange(2000)]
    u = np.concatenate(us)
    v = np.concatenate(vs)
    key = np.unique(np.concatenate([u * nv + v, v * nv + u]))
    s, t = np.divmod(key, nv)
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, s + 1, 1)
    off = np.cumsum(off).astype(np.int32)
    deg = np.diff(off)
    assert deg[1] == 4096 and deg[2] == 4097 and deg.max() == 200_001 and (deg > 4096).sum() > 2048 + 2
    sources = [0, 3, 2, 1, int(first_leaf[3]), S, int(first_leaf[4]) + 7, int(first_leaf[2]) + 123, 6, H + 1]
    return off, t.astype(np.int32), sources


@pytest.fixture(scope="module")
def hub_world(gcq, orc):
    from gcq_amd import bfs
    off, idx, sources = _hub_graph()
    print(f"[hubs] graph of {len(off) - 1} vertices, {len(idx)} edges", flush=True)
    g = bfs.Graph.from_csr(off, idx)
    refs = {s: orc.bfs_serial(off, idx, s) for s in sources}
    yield g, (off, idx), sources, refs
    g.close()


@pytest.mark.parametrize("mode", [1, 3], ids=["persistent", "bitmap"])
@pytest.mark.parametrize("threads", [4096, 65536])
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["gwfq", "glfq", "wfq", "sfq"])
def test_single_hubs(gcq, hub_world, kind, threads, mode):
    g, (off, idx), sources, refs = hub_world
    print(f"[hubs single] kind={kind} threads={threads} mode={mode}", flush=True)
    for s in sources[:6]:
        ref, reached, es = refs[s]
        d, r = g.bfs(s, kind=kind, threads=threads, block=256, mode=mode)
        bad = np.flatnonzero(d != ref)
        assert len(bad) == 0, f"source {s}: {len(bad)} depths differ, first {bad[:5]}"
        assert r["reached"] == reached and r["edges_scanned"] == es and r["levels"] == int(ref.max()) + 1


@pytest.mark.parametrize("ms", [0, 1], ids=["per_query", "multi_source"])
@pytest.mark.parametrize("threads", [4096, 65536])
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["gwfq", "glfq", "wfq", "sfq"])
def test_batch_hubs(gcq, hub_world, kind, threads, ms):
    from gcq_amd import bfs
    g, (off, idx), sources, refs = hub_world
    print(f"[hubs batch] kind={kind} threads={threads} multi_source={ms}", flush=True)
    if not ms:
        # per-query items take a frontier of the queries' summed vertices (G-WFQ rings stop at 2^25):
        # the centre (2,100 hubs at level 1), the second-level hub and the 4,097-edge hub
        sources = [sources[0], sources[5], sources[2]]
    with bfs.Batch([g]) as b:
        for use_graph in (0, 1, 1):
            outs, per, res = b.run([0] * len(sources), sources, depths=True, kind=kind, threads=threads,
                                   block=256, use_graph=use_graph, multi_source=ms)
            for i, s in enumerate(sources):
                ref, reached, es = refs[s]
                bad = np.flatnonzero(outs[i] != ref)
                assert len(bad) == 0, f"source {s}: {len(bad)} depths differ, first {bad[:5]}"
                assert per[i]["reached"] == reached and per[i]["edges_scanned"] == es
                assert per[i]["levels"] == int(ref.max()) + 1 and per[i]["status"] == 0
            assert res["levels"] == max(p["levels"] for p in per)


@pytest.mark.parametrize("ms", [0, 1], ids=["per_query", "multi_source"])
@pytest.mark.parametrize("kind", [1, 3], ids=["glfq", "sfq"])
def test_served_hubs(gcq, orc, hub_world, kind, ms):
    # The hub graph beside a small R-MAT graph in the other lane; queries
    # arrive while others run, so members join the hub graph's group at
    # different levels.
    from gcq_amd import bfs, rt, serve
    g, (off, idx), sources, refs = hub_world
    small = bfs.Graph.rmat(10, 16, 0.57, 0.19, 0.19, 11)
    scene = rt.Scene.preset(rt.RANDOM, 8, 3)
    rng = np.random.default_rng(kind)
    t = np.cumsum(rng.exponential(0.5, 40))
    rows = []
    for i in range(40):
        gi = 0 if i % 4 else 1
        src = sources[i % len(sources)] if gi == 0 else int(rng.integers(0, 1024))
        rows.append(dict(type=0, graph=gi, src=src, keep=1, id=i, arrival_ms=float(t[i])))
    print(f"[hubs serve] kind={kind} multi_source={ms}", flush=True)
    with serve.Server([g, small], scene, (1920, 1080), kind=kind, budget_bytes=1 << 30, keep_bytes=2 << 30,
                      bfs_slots=64, bfs_threads=16384, rt_threads=4096, frontier_log2=24, small_log2=11,
                      small_threads=4096, multi_source=ms) as srv:
        comps = srv.run_trace(rows)
        soff, sidx = small.csr()
        for r, c in zip(rows, comps):
            assert c["status"] == 0 and c["id"] == r["id"]
            if r["graph"] == 0:
                ref = refs[r["src"]][0]
                d = srv.fetch_depths(r["id"], len(off) - 1)
            else:
                ref = orc.bfs_serial(soff, sidx, r["src"])[0]
                d = srv.fetch_depths(r["id"], len(soff) - 1)
            bad = np.flatnonzero(d != ref)
            assert len(bad) == 0, f"request {r['id']} (graph {r['graph']}, src {r['src']}): {len(bad)} depths differ"
            assert c["levels"] == int(ref.max()) + 1
    scene.close()
    small.close()
