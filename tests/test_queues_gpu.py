"""GPU parity for the four queues through the C ABI: every emitted history
must be FIFO-linearizable under the oracle checker (the model of the
reference's Porcupine checker), the FIFO self-check must verify, and the
throughput harness must conserve items."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = [0, 1, 2, 3]  # G-WFQ, G-LFQ, WFQ, SFQ
NAMES = {0: "gwfq", 1: "glfq", 2: "wfq", 3: "sfq"}


def _queue(gcq, kind, threads, capacity=32768, **kw):
    return gcq.Queue(kind, capacity=capacity, max_threads=max(threads, 64), **kw)


def step_bound(threads, patience, help_delay):
    """Largest step count (loop iterations of one operation's fast tries,
    helping and slow path) accepted under the delay adversary below.

    Model: an operation makes its `patience` fast tries, helps at most one
    peer and then runs its own slow path.  A slow-path attempt at one ticket
    fails only when another operation's ticket lands on its slot or cell
    first, and a wave's 64 lanes draw their tickets in one aggregated FAA, so
    each concurrent wave-operation defeats at most one attempt.  A wave looks
    at announced requests every `help_delay` of its operations (G-WFQ: one
    record per lane, with the W waves' cursors staggered so that one round
    covers every thread; WFQ, YMC: every operation, help_delay = 1), after
    which it works on the request rather than against it: at most
    help_delay x W wave-operations compete with a request, W = threads / 64.
    An attempt costs at most 4 counted iterations (the counter increment and
    the slot or cell update, each retried once when a helper of the same
    request wins), paid for the operation itself and for the peer it helps.
    Measured on MI355X at 4,096 threads (W = 64): G-WFQ (help_delay 4)
    p50/p99/max 2/102/813-919 steps against a bound of 2,058; WFQ
    1/50/208-211 against 522."""
    w = threads // 64
    return patience + 2 * 4 * (1 + help_delay * w)


def _check(orc, h, what):
    rc, msg = orc.lincheck(h, orc.AUTO, 20.0)
    assert rc == orc.OK, f"{what}: {msg}"


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_config1_history_64x16_alternating(gcq, orc, kind):
    # 64 threads (block 64), 16 ops/thread alternating enq/deq, capacity 32768.
    print(f"[hist cfg1] kind={NAMES[kind]}", flush=True)
    with _queue(gcq, kind, 64) as q:
        h = q.hist_run(threads=64, ops=16, block=64, mode=0, seed=1)
    assert len(h) == 64 * 16 * 2
    enq = h[h["op"] == 0]
    assert np.all(enq["ret"] == 1) or kind == 3  # SFQ's non-blocking enqueue may report BUSY
    vals = enq["arg"]
    assert len(np.unique(vals)) == len(vals)
    _check(orc, h, f"{NAMES[kind]} cfg1")


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_split_history(gcq, orc, kind):
    print(f"[hist split] kind={NAMES[kind]}", flush=True)
    with _queue(gcq, kind, 1024) as q:
        h = q.hist_run(threads=1024, ops=16, block=256, mode=1, producer_percent=25)
    _check(orc, h, f"{NAMES[kind]} split")


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_history_65536_threads(gcq, orc, kind):
    # All four queues at 65,536 threads, the largest throughput configuration; capacity 2^20.
    print(f"[hist 64k] kind={NAMES[kind]}", flush=True)
    with gcq.Queue(kind, capacity=1 << 20, max_threads=65536) as q:
        h = q.hist_run(threads=65536, ops=4, block=256, mode=0)
    assert len(h) == 65536 * 4 * 2
    _check(orc, h, f"{NAMES[kind]} 65536 threads")


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_fifo_check_produce_then_drain(gcq, kind):
    print(f"[fifo] kind={NAMES[kind]}", flush=True)
    with _queue(gcq, kind, 4096, capacity=1 << 16) as q:
        r = q.fifo_check(threads=4096, ops=16, block=256, concurrent=False)
    assert r["verified"], r


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
def test_fifo_check_concurrent_65536(gcq, kind):
    # 25% producers x 32 ops drained concurrently, capacity 2^20.
    print(f"[fifo concurrent] kind={NAMES[kind]}", flush=True)
    with gcq.Queue(kind, capacity=1 << 20, max_threads=65536) as q:
        r = q.fifo_check(threads=65536, ops=32, block=256, concurrent=True)
    assert r["in_count"] == 16384 * 32
    assert r["verified"], r


def _conserved(q, r, what):
    """Every item enqueued over the run (warmup, probe and timed launches) was
    dequeued in it or is still in the queue: drain the quiescent queue and
    count exactly."""
    left = q.drain(4096)
    enq = r["pre_enq_success"] + r["enq_success"]
    deq = r["pre_deq_success"] + r["deq_success"]
    print(f"[{what}] enq {enq} deq {deq} left {left}", flush=True)
    assert enq == deq + left, (enq, deq, left)
    assert q.drain(64) == 0


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
@pytest.mark.parametrize("solo", [False, True], ids=["combined", "solo"])
@pytest.mark.parametrize("mode", [1, 0], ids=["split25", "balanced"])
def test_throughput_conserves_items(gcq, kind, solo, mode):
    # 8,192 threads: the smallest grid in which the harness combines claims
    print(f"[tb] kind={NAMES[kind]} solo={solo} mode={mode}", flush=True)
    with _queue(gcq, kind, 8192, capacity=1 << 16) as q:
        r = q.tb_run(mode, threads=8192, chunk_ops=32, producer_percent=25, warmup_launches=3, timed_launches=20,
                     solo_claims=solo)
        assert r["launches"] == 20 and r["elapsed_ms"] > 0
        assert r["enq_success"] > 0 and r["deq_success"] > 0
        _conserved(q, r, f"tb {NAMES[kind]}")


@pytest.mark.parametrize("kind", [0, 1, 2], ids=["gwfq", "glfq", "wfq"])
@pytest.mark.parametrize("mode", [1, 0], ids=["split25", "balanced"])
def test_grouped_tail_tickets_conserve_items(gcq, kind, mode):
    # 32,768 threads: the smallest grid in which the throughput kernel draws
    # producers' Tail tickets per workgroup (the headline kernel's form).
    print(f"[tb grouped] kind={NAMES[kind]} mode={mode}", flush=True)
    with gcq.Queue(kind, capacity=1 << 20, max_threads=32768, poll_backoff=0) as q:
        r = q.tb_run(mode, threads=32768, chunk_ops=32, producer_percent=25, warmup_launches=2, timed_launches=5)
        print(f"[tb grouped] succ {r['succ_mops']:.1f} Mops/s, enq {r['enq_success']} deq {r['deq_success']}",
              flush=True)
        assert r["enq_fail"] == 0
        assert r["enq_success"] == (32768 // 4 if mode == 1 else 32768) * 32 * 5
        assert r["deq_success"] > 0
        _conserved(q, r, f"tb grouped {NAMES[kind]}")


@pytest.mark.parametrize("kind", [0, 1, 3], ids=["gwfq", "glfq", "sfq"])
@pytest.mark.parametrize("pp", [25, 75])
def test_combined_claims_split_histories(gcq, orc, kind, pp):
    # 8,192 threads in blocks of 256 (4 waves), split roles: idle consumer
    # waves combine their claims; every EMPTY they return must still be a
    # linearizable answer for each wave of the group.
    print(f"[hist combined] kind={NAMES[kind]} pp={pp}", flush=True)
    with _queue(gcq, kind, 8192, capacity=8192) as q:
        h = q.hist_run(threads=8192, ops=8, block=256, mode=1, producer_percent=pp)
    assert len(h) == 8192 * 8
    deq = h[h["op"] == 1]
    assert np.any(deq["ret"] != 0)
    if pp == 25:
        assert np.any(deq["ret"] == 0)  # idle consumers answered EMPTY
    _check(orc, h, f"{NAMES[kind]} combined split{pp}")


def test_gwfq_slow_path_history(gcq, orc):
    # Patience 1 sends every operation through the wait-free slow path
    # (cooperative slow_faa, try_{enq,deq}_slow, helping).
    print("[hist slow path] gwfq", flush=True)
    with gcq.Queue(0, capacity=1024, max_threads=256, enq_patience=1, deq_patience=1, help_delay=4) as q:
        h = q.hist_run(threads=256, ops=16, block=64, mode=0)
    assert h["slow"].sum() > 0
    _check(orc, h, "gwfq slow path")


@pytest.mark.parametrize("kind", [0, 2], ids=["gwfq", "wfq"])
def test_wait_free_under_delay_adversary(gcq, orc, kind):
    # The two wait-free queues with an adversary: one thread in 320 (so one
    # lane in every fifth wave, which holds up its whole wave) stalls about
    # 50,000 cycles between taking a ticket and touching its slot or cell,
    # while the other waves run on and meet the stalled tickets; a patience
    # of 2 sends operations that lose a slot to the helping slow path at
    # once.  Every operation of every thread completes, the history is
    # linearizable, and the step count of each operation stays bounded.
    print(f"[adversary] kind={NAMES[kind]}", flush=True)
    kw = dict(enq_patience=2, deq_patience=2, help_delay=4) if kind == 0 else dict(enq_patience=2)
    with gcq.Queue(kind, capacity=8192, max_threads=4096, **kw) as q:
        h = q.hist_run(threads=4096, ops=16, block=256, mode=0, delay_every=320, delay_cycles=50000, seed=3)
    assert len(h) == 4096 * 16 * 2
    _check(orc, h, f"{NAMES[kind]} adversary")
    steps = h["steps"]
    print(f"[adversary] slow ops {int(h['slow'].sum())}, steps p50/p99/max "
          f"{np.percentile(steps, 50):.0f}/{np.percentile(steps, 99):.0f}/{steps.max()}", flush=True)
    assert h["slow"].sum() > 0
    bound = step_bound(4096, 2, 4 if kind == 0 else 1)
    assert steps.max() <= bound, f"max steps {steps.max()} above the modelled bound {bound}"


def test_gwfq_slow_path_past_128_laps(gcq, orc):
    # G-WFQ notes are 8-bit cycles compared modulo 256: a 64-item ring (128
    # slots) driven through about 500 laps with every operation on the slow
    # path.  Every operation completes in a few steps and the history is
    # linearizable (a note left 128 laps behind would turn slow-path enqueues
    # away from every slot for the next 128 laps).
    print("[gwfq laps] capacity 64, 64 threads, 1000 ops, patience 1", flush=True)
    with gcq.Queue(0, capacity=64, max_threads=64, enq_patience=1, deq_patience=1, help_delay=4) as q:
        h = q.hist_run(threads=64, ops=1000, block=64, mode=0)
    assert len(h) == 64 * 1000 * 2
    assert h["slow"].sum() > 0
    print(f"[gwfq laps] max steps {h['steps'].max()}", flush=True)
    assert h["steps"].max() < 2000
    _check(orc, h, "gwfq past 128 laps")


@pytest.mark.parametrize("kind", [0, 1], ids=["gwfq", "glfq"])
def test_full_queue_keeps_moving(gcq, kind):
    # Split 50/50 and 75/25 at 16,384 threads on a 65,536-item ring keep the
    # queue near full.  Producers then find the free-index ring empty; they
    # must not take plain tickets there (marking its slots ahead of the
    # consumers returning indices), which once stretched a 128-op launch from
    # ~3 ms to ~130 ms.  Items are conserved and launches stay short.
    for pp in (50, 75):
        print(f"[full queue] kind={NAMES[kind]} pp={pp}", flush=True)
        with gcq.Queue(kind, capacity=65536, max_threads=65536, poll_backoff=0) as q:
            r = q.tb_run(1, threads=16384, chunk_ops=128, producer_percent=pp, warmup_launches=5, timed_launches=20)
            print(f"[full queue] succ {r['succ_mops']:.1f} Mops/s, launch {r['kernel_ms_mean']:.2f} ms", flush=True)
            _conserved(q, r, f"full queue {NAMES[kind]} pp={pp}")
        assert r["kernel_ms_mean"] < 25.0


@pytest.mark.parametrize("kind", [0, 1], ids=["gwfq", "glfq"])
def test_remap_history_and_conservation(gcq, orc, kind):
    # remap = 1 spreads consecutive tickets over 128-B lines (cache_remap,
    # gwf_queue.hpp:218-230, glf_queue.hpp:116-120).  A 1,024-item ring under
    # 1,024 threads x 16 ops turns many laps through the permuted slots.
    print(f"[remap] kind={NAMES[kind]}", flush=True)
    with _queue(gcq, kind, 1024, capacity=1024, remap=1) as q:
        h = q.hist_run(threads=1024, ops=16, block=256, mode=0)
    _check(orc, h, f"{NAMES[kind]} remap")
    with gcq.Queue(kind, capacity=1 << 16, max_threads=8192, remap=1) as q:
        r = q.tb_run(1, threads=8192, chunk_ops=32, producer_percent=25, warmup_launches=3, timed_launches=20)
        assert r["enq_success"] > 0 and r["deq_success"] > 0
        _conserved(q, r, f"remap {NAMES[kind]}")


@pytest.mark.parametrize("mode", [1, 0], ids=["split25", "balanced"])
def test_wfq_bounded_pool_ms_window(gcq, orc, mode):
    # WFQ on its default pool at capacity 2^20 and 65,536 threads: a ring of
    # 2^22 cells (64 MiB), reclaimed while the queue runs.  128 ops per
    # launch, a 200 ms warmup and a 1,000 ms timed window: about 2 * 10^9
    # cell tickets through 2^22 cells.  No enqueue may be refused, and every item
    # is accounted for.
    print(f"[wfq ring] mode={mode}", flush=True)
    with gcq.Queue(2, capacity=1 << 20, max_threads=65536) as q:
        assert q.cfg.wfq_segment_cells == 0  # the library's default pool
        r = q.tb_run(mode, threads=65536, chunk_ops=128, producer_percent=25, warmup_ms=200, run_ms=1000)
        print(f"[wfq ring] {r['succ_mops']:.1f} Mops/s, launches {r['launches']}", flush=True)
        assert r["enq_fail"] == 0
        assert r["succ_mops"] > 100
        _conserved(q, r, f"wfq ring mode={mode}")
        # the same queue, after ~10^9 tickets, still gives linearizable histories
        h = q.hist_run(threads=65536, ops=4, block=256, mode=0)
    assert len(h) == 65536 * 4 * 2
    _check(orc, h, "wfq ring after a long run")
