"""The oracle FIFO linearizability checker, pinned against the reference's
golden histories, hand-built illegal histories, and exact-vs-pattern
agreement on random small histories."""
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN

FIXTURES = ["history_glfq_m0_t64_o16.jsonl", "history_gwfq_m0_t32_o8.jsonl"]


def _ops(rows):
    from oracle.pyoracle import OP_DTYPE
    return np.array([tuple(r) for r in rows], dtype=OP_DTYPE)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("engine", [1, 2])
def test_reference_fixtures_accepted(orc, name, engine):
    ops = orc.read_jsonl(os.path.join(GOLDEN, name))
    assert len(ops) in (2048, 512)
    rc, msg = orc.lincheck(ops, engine)
    assert rc == orc.OK, msg


def _mutants(ops):
    """Illegal variants of a legal alternating history."""
    deq = np.flatnonzero((ops["op"] == 1) & (ops["ret"] != 0))
    enq = {int(v): i for i, v in enumerate(ops["arg"]) if ops["op"][i] == 0}
    out = {}
    m = ops.copy()
    m["ret"][deq[5]] = 0xDEADBEEF  # value never enqueued
    out["fresh"] = m
    m = ops.copy()
    m["ret"][deq[7]] = m["ret"][deq[3]]  # same value dequeued twice
    out["repeat"] = m
    # swap the results of two dequeues that are ordered in real time and whose
    # enqueues are ordered in real time
    done = False
    for a in deq:
        for b in deq:
            if done:
                break
            va, vb = int(ops["ret"][a]), int(ops["ret"][b])
            ea, eb = enq[va], enq[vb]
            if ops["end"][a] < ops["call"][b] and ops["end"][ea] < ops["call"][eb]:
                m = ops.copy()
                m["ret"][a], m["ret"][b] = vb, va
                out["order"] = m
                done = True
    # a dequeue made "empty" right after a completed enqueue with no concurrent dequeue
    m = ops.copy()
    m["ret"][deq[len(deq) // 2]] = 0
    out["witness_or_lost"] = m
    return out


@pytest.mark.parametrize("name", FIXTURES)
def test_mutants_rejected(orc, name):
    ops = orc.read_jsonl(os.path.join(GOLDEN, name))
    muts = _mutants(ops)
    assert set(muts) == {"fresh", "repeat", "order", "witness_or_lost"}
    for kind, m in muts.items():
        rc, msg = orc.lincheck(m, orc.PATTERN)
        assert rc == orc.ILLEGAL, (kind, msg)
        # The exact search may exhaust its budget proving a 2,048-op history
        # illegal (as Porcupine can); it must never accept one.
        rc, msg = orc.lincheck(m, orc.WGL, 3.0)
        assert rc in (orc.ILLEGAL, orc.UNKNOWN), (kind, msg)


def test_empty_witness_needs_union_of_intervals(orc):
    # v1 surely present on (2,5), v2 surely present on (4.5,inf): together they
    # cover the empty dequeue [3,9] although neither does alone.
    rows = [(0, 0, 11, 1, 1, 2),     # enq v1 [1,2]
            (1, 0, 22, 1, 3, 4),     # enq v2 [3,4]
            (2, 1, 0, 0, 5, 12),     # deq -> empty [5,12]
            (3, 1, 0, 11, 6, 7)]     # deq -> v1 [6,7]
    for engine in (1, 2):
        assert orc.lincheck(_ops(rows), engine)[0] == orc.ILLEGAL
    # the same history with v2's enqueue overlapping the dequeue of v1's gap is legal
    rows = [(0, 0, 11, 1, 1, 2), (3, 1, 0, 11, 3, 5), (2, 1, 0, 0, 4, 9), (1, 0, 22, 1, 6, 8)]
    for engine in (1, 2):
        assert orc.lincheck(_ops(rows), engine)[0] == orc.OK


def test_failed_enqueue_is_a_noop(orc):
    rows = [(0, 0, 5, 0, 1, 2), (1, 1, 0, 0, 3, 4)]  # failed enq, then empty deq
    assert orc.lincheck(_ops(rows), orc.WGL)[0] == orc.OK
    assert orc.lincheck(_ops(rows), orc.PATTERN)[0] == orc.OK


def _random_history(rng, nproc, nops, corrupt):
    """A history produced by a sequential FIFO at random linearization points
    inside random intervals; optionally corrupt one result."""
    events = []
    t = 0
    ops = []
    val = 1
    pending = []  # (proc, kind, value)
    for p in range(nproc):
        for k in range(nops):
            if rng.random() < 0.55:
                pending.append((p, 0, val))
                val += 1
            else:
                pending.append((p, 1, 0))
    rng.shuffle(pending)
    # per-process sequential: assign intervals in order per proc
    clock = [0] * nproc
    timeline = []
    for (p, kind, v) in pending:
        start = clock[p] + rng.randint(1, 4)
        lin = start + rng.random() * 3
        end = lin + rng.random() * 3
        clock[p] = int(end) + 1
        timeline.append((lin, p, kind, v, start, end))
    timeline.sort()
    q = []
    rows = []
    for (lin, p, kind, v, start, end) in timeline:
        if kind == 0:
            q.append(v)
            rows.append([p, 0, v, 1, start, end])
        else:
            r = q.pop(0) if q else 0
            rows.append([p, 1, 0, r, start, end])
    # make all timestamps distinct integers preserving order
    stamps = sorted({x for r in rows for x in (r[4], r[5])})
    rank = {s: i + 1 for i, s in enumerate(stamps)}
    for r in rows:
        r[4], r[5] = rank[r[4]], rank[r[5]]
    fix = []
    for r in rows:
        if r[4] == r[5]:
            r[5] = r[4]  # cannot happen after ranking distinct floats
        fix.append(r)
    # de-duplicate equal stamps by small offsets: rebuild with 2 slots per op
    evs = []
    for i, r in enumerate(fix):
        evs.append((r[4], 0, i))
        evs.append((r[5], 1, i))
    evs.sort()
    for k, (_, which, i) in enumerate(evs):
        fix[i][4 + which] = k + 1
    if corrupt:
        i = rng.randrange(len(fix))
        if fix[i][1] == 1:
            choices = [0] + [r[2] for r in fix if r[1] == 0]
            fix[i][3] = rng.choice(choices)
    return _ops([tuple(r) for r in fix])


def test_engines_agree_on_random_histories(orc):
    rng = random.Random(1234)
    n_illegal = 0
    for it in range(600):
        h = _random_history(rng, nproc=rng.randint(2, 4), nops=rng.randint(1, 4), corrupt=it % 2 == 1)
        a, ma = orc.lincheck(h, orc.WGL, 5.0)
        b, mb = orc.lincheck(h, orc.PATTERN)
        assert a != orc.UNKNOWN
        assert a == b, (it, a, ma, b, mb, h)
        n_illegal += a == orc.ILLEGAL
        if it % 2 == 0:
            assert a == orc.OK, (it, ma)
    assert n_illegal > 50  # the corrupted half really exercises the rejecting paths
