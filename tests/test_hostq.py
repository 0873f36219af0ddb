"""The oracle's host std::atomic restatements of the four queues (the CPU
baseline of the throughput line): their histories must pass the same FIFO
linearizability check as the GPU's, and the timed driver must conserve
items.  CPU only."""
import numpy as np
import pytest

KINDS = [0, 1, 2, 3]
NAMES = {0: "gwfq", 1: "glfq", 2: "wfq", 3: "sfq"}


def _check(orc, h, what):
    rc, msg = orc.lincheck(h, orc.AUTO, 20.0)
    assert rc == orc.OK, f"{what}: {msg}"


@pytest.mark.parametrize("kind", KINDS, ids=[NAMES[k] for k in KINDS])
@pytest.mark.parametrize("mode", [0, 1], ids=["alternating", "split25"])
def test_host_history_linearizable(orc, kind, mode):
    h, _ = orc.hist_run(kind, threads=8, ops=64, mode=mode, producer_percent=25, capacity=64)
    assert len(h) == (8 * 64 * 2 if mode == 0 else 8 * 64)
    enq = h[(h["op"] == 0) & (h["ret"] == 1)]
    assert len(np.unique(enq["arg"])) == len(enq)
    _check(orc, h, f"host {NAMES[kind]} mode {mode}")


def test_host_gwfq_slow_path_under_delays(orc):
    # Patience 1 routes every operation through the wait-free slow path;
    # every fourth thread sleeps between taking a ticket and using it.
    h, steps = orc.hist_run(0, threads=8, ops=32, mode=0, capacity=16, enq_patience=1, deq_patience=1,
                            delay_every=4, delay_us=50)
    _check(orc, h, "host gwfq slow path")
    assert 0 < steps < 10_000


def test_host_tb_conserves_items(orc):
    for kind in KINDS:
        r = orc.tb_run(kind, threads=4, mode=1, producer_percent=50, chunk_ops=32, warmup_ms=5, run_ms=40,
                       capacity=1024)
        assert r["enq_success"] > 0 and r["deq_success"] > 0, NAMES[kind]
        # what was dequeued in the window was enqueued in it or queued before it
        assert r["deq_success"] <= r["enq_success"] + 1024, NAMES[kind]
        assert r["succ_mops"] > 0


def test_host_gwfq_notes_stay_fresh_past_128_laps(orc):
    # A 16-item ring (32 slots) driven through about 375 laps with every
    # operation on the slow path (patience 1).  Notes are 8-bit cycles
    # compared modulo 256.  Under the reference's rule the fast path and the
    # dequeuers' marks keep an entry's old note, so once the ring has turned
    # 128 times the initial notes (0xFF) read as ahead of every ticket and a
    # slow-path enqueue is turned away by every slot until the ring turns
    # another 128 times: thousands of steps for one operation.  With the
    # library's rule (a note behind the written cycle becomes cycle - 1) every
    # operation stays within a few steps.
    kw = dict(threads=2, ops=6000, mode=0, capacity=16, enq_patience=1, deq_patience=1)
    ref_steps = max(orc.hist_run(0, ref_slow_rule=True, **kw)[1] for _ in range(3))
    h, steps = orc.hist_run(0, **kw)
    assert ref_steps > 32 * 64
    assert steps < 100
    _check(orc, h, "host gwfq past 128 laps")
