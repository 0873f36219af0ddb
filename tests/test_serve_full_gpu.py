"""The served mix at its benchmark scale: resident R-MAT graphs of scales
14..22 (edge factor 16, seed 11), the 1920x1080 64-sphere scene, a 64 GB
request budget and the benchmark's server configuration (a 2^28-slot
frontier ring, 245,760 BFS + 16,384 RT threads in blocks of 512), with one
BFS lane and with 32,768 of the BFS threads in a second lane for the graphs
of at most 2^18 vertices, with and without bit-parallel groups for the
queries on one graph.  About 60
requests, 90% BFS queries on sources in each graph's giant component, 10%
256 x 256 tiles, served both as a Poisson trace at 150 requests/s and as one
burst.  Every BFS depth array equals the serial BFS oracle's bit for bit;
every tile equals the same tile rendered alone, and the CPU restatement of
the renderer within 1e-5."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

FRAME = (1920, 1080)
SCALES = list(range(14, 23))
SERVER = dict(bfs_threads=245760, block=512, rt_threads=16384, frontier_log2=28, budget_bytes=64 << 30,
              keep_bytes=4 << 30)
LANES = {1: {}, 2: dict(small_log2=18, small_threads=32768),
         "ms": dict(small_log2=18, small_threads=32768, multi_source=1)}


@pytest.fixture(scope="module")
def big_world(gcq, orc):
    from gcq_amd import bfs, rt
    print(f"[serve full] building R-MAT scales {SCALES[0]}..{SCALES[-1]}", flush=True)
    graphs = [bfs.Graph.rmat(s, 16, 0.57, 0.19, 0.19, 11) for s in SCALES]
    csrs = [g.csr() for g in graphs]
    scene = rt.Scene.preset(rt.RANDOM, 64, 3)
    full, _ = orc.rt_render(*scene.arrays(), FRAME[0], FRAME[1])
    print("[serve full] graphs and reference frame ready", flush=True)
    yield graphs, scene, csrs, full
    scene.close()
    for g in graphs:
        g.close()


@pytest.mark.parametrize("lanes", [1, 2, "ms"], ids=["one_lane", "two_lanes", "two_lanes_ms"])
@pytest.mark.parametrize("rate", [150.0, 1e9], ids=["poisson150", "burst"])
def test_cfg5_scale_mix_is_exact(gcq, orc, big_world, rate, lanes):
    from gcq_amd import serve
    graphs, scene, csrs, full = big_world
    nvs = [len(c[0]) - 1 for c in csrs]
    rows = serve.poisson_trace(60, rate, nvs, FRAME, rt_fraction=0.1, seed=13, keep=1,
                               graph_sources=lambda gi, i: graphs[gi].pick_source(11, i))
    nb = sum(r["type"] == 0 for r in rows)
    print(f"[serve full] rate={rate:g}: {nb} BFS, {len(rows) - nb} tiles, "
          f"scales {sorted({SCALES[r['graph']] for r in rows if r['type'] == 0})}", flush=True)
    assert any(r["type"] == 0 and SCALES[r["graph"]] == 22 for r in rows)
    assert any(r["type"] == 1 for r in rows)
    with serve.Server(graphs, scene, FRAME, **SERVER, **LANES[lanes]) as srv:
        comps = srv.run_trace(rows)
        print(f"[serve full] lanes={lanes}", serve.latency_summary(comps), flush=True)
        assert len(comps) == len(rows)
        for r, c in zip(rows, comps):
            assert c["status"] == 0 and c["id"] == r["id"] and c["type"] == r["type"]
            assert c["done_ms"] >= c["admit_ms"] >= c["arrival_ms"] - 1e-6
            if r["type"] == 0:
                off, idx = csrs[r["graph"]]
                ref, _, _ = orc.bfs_serial(off, idx, r["src"])
                d = srv.fetch_depths(r["id"], len(off) - 1)
                bad = np.flatnonzero(d != ref)
                assert len(bad) == 0, (f"request {r['id']} (scale {SCALES[r['graph']]}, src {r['src']}): "
                                       f"{len(bad)} depths differ, first at {bad[:5]}")
                assert c["levels"] == int(ref.max()) + 1
            else:
                img = srv.fetch_image(r["id"], r["w"], r["h"])
                alone, res = scene.render(width=FRAME[0], height=FRAME[1], tile_w=256, tile_h=256, x0=r["x0"],
                                          y0=r["y0"], w=r["w"], h=r["h"])
                assert np.array_equal(img, alone), f"request {r['id']}: tile differs from the tile alone"
                assert c["rays"] == res["total_rays"]
                ref = full[r["y0"]:r["y0"] + r["h"], r["x0"]:r["x0"] + r["w"]]
                assert float(np.abs(img - ref).max()) <= 1e-5
    print(f"[serve full] rate={rate:g}: {len(rows)} results exact", flush=True)
