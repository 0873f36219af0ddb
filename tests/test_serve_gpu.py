"""GPU tests of the continuously batched server: every BFS result equals
the serial BFS oracle (and a plain run) bit for bit, every tile equals the
same tile rendered alone (and the CPU restatement within 1e-5), whatever
else was in flight; admission never exceeds the slot limits."""
import numpy as np
import pytest

# A stuck server fails its test quickly (the library's own stop deadline is 30 s).
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(100)]

FRAME = (1920, 1080)


@pytest.fixture(scope="module")
def world(gcq):
    from gcq_amd import bfs, rt
    graphs = [bfs.Graph.rmat(s, 16, 0.57, 0.19, 0.19, 11) for s in (9, 11, 13)]
    graphs.append(bfs.Graph.lattice(200, 200, 0.75, 0.01, 7))
    scene = rt.Scene.preset(rt.RANDOM, 64, 3)
    csrs = [g.csr() for g in graphs]
    yield graphs, scene, csrs
    scene.close()
    for g in graphs:
        g.close()


def _check_results(orc, srv, rows, comps, graphs, scene, csrs, full_frame=None):
    assert len(comps) == len(rows)
    for r, c in zip(rows, comps):
        assert c["status"] == 0 and c["id"] == r["id"] and c["type"] == r["type"]
        assert c["done_ms"] >= c["admit_ms"] >= c["arrival_ms"] - 1e-6
        if r["type"] == 0:
            off, idx = csrs[r["graph"]]
            ref, _, _ = orc.bfs_serial(off, idx, r["src"])
            d = srv.fetch_depths(r["id"], len(off) - 1)
            assert np.array_equal(d, ref), f"request {r['id']}: depths differ"
            assert c["levels"] == int(ref.max()) + 1
        else:
            img = srv.fetch_image(r["id"], r["w"], r["h"])
            alone, res = scene.render(width=FRAME[0], height=FRAME[1], tile_w=256, tile_h=256, x0=r["x0"], y0=r["y0"],
                                      w=r["w"], h=r["h"])
            assert np.array_equal(img, alone), f"request {r['id']}: tile differs from the tile alone"
            assert c["rays"] == res["total_rays"]
            if full_frame is not None:
                ref = full_frame[r["y0"]:r["y0"] + r["h"], r["x0"]:r["x0"] + r["w"]]
                assert float(np.abs(img - ref).max()) <= 1e-5


# Two lanes: queries on graphs of at most 2^11 vertices (R-MAT scales 9 and
# 11) run in a second BFS sub-grid of 4,096 threads with its own barrier.
# "ms": two lanes, and queries on one graph share bit-parallel groups.
LANES = {1: {}, 2: dict(small_log2=11, small_threads=4096), "ms": dict(small_log2=11, small_threads=4096,
                                                                       multi_source=1)}


@pytest.mark.parametrize("lanes", [1, 2, "ms"], ids=["one_lane", "two_lanes", "two_lanes_ms"])
def test_mix_equals_alone(gcq, orc, world, lanes):
    from gcq_amd import serve
    graphs, scene, csrs = world
    nvs = [len(c[0]) - 1 for c in csrs]
    rows = serve.poisson_trace(150, 3000.0, nvs, FRAME, rt_fraction=0.1, seed=3, keep=1)
    full, _ = orc.rt_render(*scene.arrays(), FRAME[0], FRAME[1])
    print(f"[serve mix] {len(rows)} requests, {lanes} lane(s)", flush=True)
    with serve.Server(graphs, scene, FRAME, budget_bytes=1 << 30, keep_bytes=1 << 30, bfs_threads=16384,
                      rt_threads=8192, frontier_log2=22, **LANES[lanes]) as srv:
        comps = srv.run_trace(rows)
        print("[serve mix]", serve.latency_summary(comps), flush=True)
        _check_results(orc, srv, rows, comps, graphs, scene, csrs, full)


@pytest.mark.parametrize("lanes", [1, "ms"], ids=["one_lane", "two_lanes_ms"])
def test_sparse_arrivals_after_idle_parks(gcq, orc, world, lanes):
    # Requests about 15 ms apart: between them every lane is idle and parks
    # level after level; each arrival must still be admitted and exact.
    from gcq_amd import serve
    graphs, scene, csrs = world
    nvs = [len(c[0]) - 1 for c in csrs]
    rows = serve.poisson_trace(16, 66.0, nvs, FRAME, rt_fraction=0.25, seed=12, keep=1)
    print(f"[serve sparse] {len(rows)} requests over {rows[-1]['arrival_ms']:.0f} ms, {lanes} lane(s)", flush=True)
    with serve.Server(graphs, scene, FRAME, budget_bytes=1 << 30, keep_bytes=1 << 30, bfs_threads=16384,
                      rt_threads=8192, frontier_log2=22, **LANES[lanes]) as srv:
        comps = srv.run_trace(rows)
        print("[serve sparse]", serve.latency_summary(comps), flush=True)
        _check_results(orc, srv, rows, comps, graphs, scene, csrs)


@pytest.mark.parametrize("lanes", [1, 2, "ms"], ids=["one_lane", "two_lanes", "two_lanes_ms"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["gwfq", "glfq", "wfq", "sfq"])
def test_burst_with_tight_limits(gcq, orc, world, kind, lanes):
    # Everything arrives at once; 3 BFS slots, 2 tile slots and a budget of
    # about two of the largest queries force requests to wait for room
    # (kept results live in their own keep_bytes arena).
    from gcq_amd import serve
    graphs, scene, csrs = world
    nvs = [len(c[0]) - 1 for c in csrs]
    rows = serve.poisson_trace(40, 1e9, nvs, FRAME, tile=128, rt_fraction=0.25, seed=kind, keep=1)
    budget = 2 * 4 * max(nvs) + 2 * 128 * 128 * 76 + 4096
    print(f"[serve burst] kind={kind} lanes={lanes}", flush=True)
    with serve.Server(graphs, scene, FRAME, kind=kind, budget_bytes=budget, keep_bytes=1 << 28, bfs_slots=3,
                      rt_slots=2, bfs_threads=4096 if lanes == 1 else 8192, rt_threads=4096, frontier_log2=20,
                      **LANES[lanes]) as srv:
        comps = srv.run_trace(rows)
        _check_results(orc, srv, rows, comps, graphs, scene, csrs)
    # at no time more than 3 BFS queries / 2 tiles between admission and completion
    for typ, lim in ((0, 3), (1, 2)):
        ev = sorted([(c["admit_ms"], 1) for c in comps if c["type"] == typ] +
                    [(c["done_ms"], -1) for c in comps if c["type"] == typ])
        live = peak = 0
        for _, dlt in ev:
            live += dlt
            peak = max(peak, live)
        assert peak <= lim


def test_submit_poll_api_and_errors(gcq, orc, world):
    from gcq_amd import GcqError, serve
    graphs, scene, csrs = world
    with serve.Server(graphs, scene, FRAME, budget_bytes=1 << 28, keep_bytes=1 << 28, bfs_threads=4096,
                      rt_threads=4096, frontier_log2=20) as srv:
        srv.start()
        with pytest.raises(GcqError):
            srv.submit(type=0, graph=99, src=0, id=1)
        with pytest.raises(GcqError):
            srv.submit(type=1, x0=1800, y0=0, w=256, h=256, id=2)  # outside the frame
        with pytest.raises(GcqError):
            srv.submit(type=1, x0=0, y0=0, w=512, h=256, id=3)  # above 65536 pixels
        srv.submit(type=0, graph=1, src=5, id=10, keep=1)
        srv.submit(type=1, x0=0, y0=0, w=64, h=64, id=11, keep=1)
        srv.submit(type=0, graph=3, src=10**9, id=12, keep=1)  # out of range -> vertex 0
        got = []
        import time
        t = time.time()
        while len(got) < 3 and time.time() - t < 60:
            got += srv.poll()
        assert sorted(c["id"] for c in got) == [10, 11, 12]
        srv.stop()
        off, idx = csrs[1]
        assert np.array_equal(srv.fetch_depths(10, len(off) - 1), orc.bfs_serial(off, idx, 5)[0])
        off, idx = csrs[3]
        assert np.array_equal(srv.fetch_depths(12, len(off) - 1), orc.bfs_serial(off, idx, 0)[0])
        alone, _ = scene.render(width=FRAME[0], height=FRAME[1], tile_w=256, tile_h=256, x0=0, y0=0, w=64, h=64)
        assert np.array_equal(srv.fetch_image(11, 64, 64), alone)


def test_plain_trace_matches_levels(gcq, orc, world):
    # Served without kept results: depth arrays come from the budget arena
    # and are reused by later requests.
    from gcq_amd import serve
    graphs, scene, csrs = world
    nvs = [len(c[0]) - 1 for c in csrs]
    rows = serve.poisson_trace(30, 2000.0, nvs, FRAME, rt_fraction=0.2, seed=9)
    plain = serve.plain_run_trace(graphs, scene, FRAME, rows)
    with serve.Server(graphs, scene, FRAME, budget_bytes=1 << 28, bfs_threads=4096, rt_threads=4096,
                      frontier_log2=20) as srv:
        served = srv.run_trace(rows)
    for r, a, b in zip(rows, plain, served):
        want = 0
        if r["type"] == 0:
            off, idx = csrs[r["graph"]]
            want = int(orc.bfs_serial(off, idx, r["src"])[0].max()) + 1
        detail = f"request {r} plain {a} served {b} oracle levels {want}"
        assert a["levels"] == want, detail
        assert b["levels"] == want, detail
        assert a["rays"] == b["rays"], detail
