"""CPU model of how the grid shares a level's long adjacencies
(ms_hubs in gcq_amd/csrc/bfs_batch.hpp): posted vertices take consecutive
chunk ranges (one 64-bit CAS adds an entry and its chunk count, so entry k's
first chunk is the sum of the counts before it); wave w < P, P = min(chunks,
waves), scans chunk w, then chunks P + s + 8k drawn by fetch-add on cursor s,
moving on to cursor s + 1 (mod 8) when one runs out; a chunk maps to the
last entry whose first chunk is at or below it.  Under random interleavings
of the waves, every edge of every posted vertex must be scanned exactly once,
and a wave must stop after at most 8 failed draws."""
import numpy as np
import pytest

CHUNK, CURSORS = 1024, 8


def _post(degs):
    bases, total = [], 0
    for d in degs:
        bases.append(total)
        total += -(-d // CHUNK)
    return bases, total


def _entry_of(bases, c):
    lo, hi = 0, len(bases)
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if bases[mid] <= c:
            lo = mid
        else:
            hi = mid
    return lo


def _run(degs, nwaves, rng, stride=CURSORS):
    bases, total = _post(degs)
    P = min(total, nwaves)
    cursor = [0] * CURSORS
    scanned = [np.zeros(d, np.int64) for d in degs]
    # per wave: (next chunk or None, cursor index, cursors tried)
    state = {w: [w, w % CURSORS, 0] for w in range(P)}
    draws = {w: 0 for w in range(P)}
    while state:
        w = int(rng.choice(list(state)))
        c, s, tried = state[w]
        if c is None or c >= total:
            k = cursor[s]
            cursor[s] += 1
            c = P + s + stride * k
            draws[w] += 1
            if c >= total:
                s, tried = (s + 1) % CURSORS, tried + 1
                state[w] = [None, s, tried]
                if tried == CURSORS:
                    del state[w]
                continue
        e = _entry_of(bases, c)
        j = c - bases[e]
        first = j * CHUNK
        scanned[e][first:min(first + CHUNK, degs[e])] += 1
        state[w] = [None, s, tried]
    return scanned, draws, total


@pytest.mark.parametrize("degs,nwaves", [
    ([4097], 64),                      # one chunk of one edge at the end
    ([200_001], 64),                   # more chunks than waves
    ([200_001], 4096),                 # more waves than chunks
    ([4097, 4098, 200_001, 5000], 8),  # fewer waves than cursors
    ([10_000] * 300, 1000),
    ([4097 + i for i in range(2048)], 3712),
], ids=["one", "big_few_waves", "big_many_waves", "waves_below_cursors", "many", "full_table"])
def test_every_edge_scanned_once(degs, nwaves):
    rng = np.random.default_rng(len(degs) + nwaves)
    scanned, draws, total = _run(degs, nwaves, rng)
    for d, sc in zip(degs, scanned):
        assert len(sc) == d and (sc == 1).all()
    # each wave draws its share plus at most one failed draw per cursor
    assert sum(draws.values()) <= max(0, total - min(total, nwaves)) + CURSORS * min(total, nwaves)


def test_planted_wrong_stride_is_caught():
    # A stride of 7 instead of 8 on the striped cursors scans some chunks
    # twice; the model's check must see it.
    scanned, _, _ = _run([200_001], 64, np.random.default_rng(1), stride=7)
    assert not (scanned[0] == 1).all()
