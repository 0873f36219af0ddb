// bfs_batch.hpp — the multi-query BFS level expansion shared by the batched
// (batch.hip) and served (serve.hip) persistent kernels.
//
// Work items are (query slot, vertex) packed in 32 bits: the vertex in the
// low `vbits` bits, the slot above it, below 2^31 so no queue sentinel can
// occur.  A vertex discovered while expanding level L of a query that
// started at level `start` gets depth L - start + 1 through a CAS on -1
// (bfs_bench.cpp:383-406), so each query's depths are exactly BFS depths.
#pragma once
#include "frontier.hpp"

namespace bk {

using namespace gcqf;

// One query, read by the kernel with device-scope loads.
struct Slot {
  uint64_t off, idx, depth;  // CSR of its graph, its depth array
  uint32_t start, src;       // level at which it starts; its source
  uint32_t nv, pad;
  uint64_t visited;          // optional bitmap, one bit per vertex (0 = none)
  uint64_t pad2[2];
};
static_assert(sizeof(Slot) == 64, "slot");

constexpr uint32_t kAliveStride = 32;  // one 128-B line per query

struct Dev {
  const Slot* slots;
  uint32_t* alive;  // [slot * kAliveStride]: last level at which the query has work
  int32_t* done;    // per slot: levels once finished, -1 while running
  uint32_t* err;    // bit 0: a frontier enqueue failed
  uint32_t nslots, vbits, pop_width, flush_at;
};

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int lane) {
  const uint32_t lo = __shfl((uint32_t)x, lane, 64);
  const uint32_t hi = __shfl((uint32_t)(x >> 32), lane, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Drain the level's input queue; returns the lane's discoveries.  All 64
// lanes of the wave call it together.  kLive: slots and depth arrays may be
// rewritten while the kernel runs (the server admits queries into freed
// slots), so they are read at device scope; otherwise (one batch per launch)
// slots are fixed and a depth only goes from -1 to a level, so the visited
// filter and the slot fields are read through the caches.
//
// The kernels are level-synchronous (a grid barrier ends every level, and
// every push of a level lands in its output queue), so the input queue is
// drained by counted pops (F::pop_counted: one FAA on Head against the
// level's end, read at its start, and no claim reads).  The caller sets a
// drained queue's Head back to its end (F::drained) and, when the queue kind
// keeps Head there (kHeadFromDrain), passes that end as h0 when the queue
// next feeds a level: a level whose tickets [h0, end) fit in the grid's
// popping lanes (grid_waves x pop_width) then takes them by grid position,
// lane l of wave w ticket h0 + l * grid_waves + w, with no FAA on Head.
// end_out receives the level's end.
template <class F, bool kLive = true>
__device__ uint32_t expand(const F& qin, const F& qout, const Dev& d, int32_t L, uint32_t tid, uint32_t* stage,
                           uint64_t h0, uint64_t grid_waves, uint64_t* end_out, HubPool* hp) {
  auto rd = [](const auto* p) { return kLive ? gcq::ld_rlx(p) : gcq::ld_cached(p); };
  const uint32_t lane = lane_id();
  const uint32_t vmask = (1u << d.vbits) - 1u;
  uint32_t sc = 0, disc_n = 0;
  const bool pops = lane < d.pop_width;
  const uint64_t tix = h0 + (uint64_t)lane * grid_waves + tid / 64;
  uint64_t first_word = 0;
  if constexpr (F::kHeadFromDrain) {
    if (h0 != kNoHead && pops) first_word = qin.peek(tix);
  }
  const uint64_t end = qin.level_end();
  *end_out = end;
  const bool one_round = F::kHeadFromDrain && h0 != kNoHead && end >= h0 && end - h0 <= grid_waves * d.pop_width;
  for (;;) {
    uint32_t item = 0;
    int st = -1;
    if (pops) {
      if constexpr (F::kHeadFromDrain) {
        st = one_round ? qin.take(tid, tix, end, first_word, &item) : qin.pop_counted(tid, end, &item);
      } else {
        st = qin.pop_counted(tid, end, &item);
      }
    }
    if (__ballot(st >= 0) == 0) break;  // every popping lane drew a ticket past the level's end
    const bool got = st > 0;
    uint32_t slot = 0;
    int32_t beg = 0, deg = 0, dn = 0;
    uint64_t idxp = 0, depp = 0, vis = 0;
    if (got) {
      slot = item >> d.vbits;
      const uint32_t v = item & vmask;
      const Slot* s = &d.slots[slot];
      const int32_t* off = (const int32_t*)rd(&s->off);
      idxp = rd(&s->idx);
      depp = rd(&s->depth);
      if (!kLive) vis = rd(&s->visited);
      dn = L - (int32_t)rd(&s->start) + 1;
      beg = off[v];  // CSR arrays are uploaded before launch and never written
      deg = off[v + 1] - beg;
      if (hp && deg > kHubDeg && hub_post(hp, item, beg, deg, 0ull)) deg = 0;  // the grid scans it
    }
    int32_t incl = deg;  // inclusive prefix of degrees across the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += t;
    }
    const int32_t excl = incl - deg;
    const int32_t total = __shfl(incl, 63, 64);
    auto owner = [&](int32_t base) {  // first lane whose inclusive prefix exceeds this lane's edge
      const int32_t e = base + (int32_t)lane;
      const int32_t ec = e < total ? e : total - 1;
      int lo = 0, hi = 63;
      for (int s = 0; s < 6; ++s) {
        const int mid = (lo + hi) >> 1;
        if (__shfl(incl, mid, 64) > ec) hi = mid;
        else lo = mid + 1;
      }
      return lo;
    };
    auto neighbour = [&](int32_t base, int lo) -> uint32_t {  // this lane's edge target, 0 past the end
      const int32_t e = base + (int32_t)lane;
      const int32_t obeg = __shfl(beg, lo, 64);
      const int32_t oexcl = __shfl(excl, lo, 64);
      const uint64_t oidx = shfl64(idxp, lo);
      return e < total ? (uint32_t)((const int32_t*)oidx)[obeg + (e - oexcl)] : 0u;
    };
    // The next round's neighbour indices are loaded before this round's
    // probes and CASes, so their latency overlaps this round's chain instead
    // of adding to it (256 R-MAT queries: 2.5% alone, 6% with the direct CAS
    // below).
    int lo_next = total > 0 ? owner(0) : 0;
    uint32_t w_next = total > 0 ? neighbour(0, lo_next) : 0u;
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t e = base + (int32_t)lane;
      const int lo = lo_next;
      const uint32_t w_here = w_next;
      if (base + 64 < total) {
        lo_next = owner(base + 64);
        w_next = neighbour(base + 64, lo_next);
      }
      const uint32_t oslot = __shfl(slot, lo, 64);
      const int32_t odn = __shfl(dn, lo, 64);
      const uint64_t odep = shfl64(depp, lo);
      const uint64_t ovis = kLive ? 0 : shfl64(vis, lo);
      bool disc = false;
      uint32_t w = 0;
      if (e < total) {
        w = w_here;
        int32_t* dp = (int32_t*)odep + w;
        // The bitmap packs 32 vertices per word where the depth array holds
        // one, so most visited probes are served by a cached line; a bit read
        // stale (clear) only sends the edge on to the CAS.  With a bitmap, a
        // clear bit goes straight to the CAS (the depth probe would find -1
        // nearly always); without one, the depth is probed first.
        uint32_t* vw = ovis ? (uint32_t*)ovis + (w >> 5) : nullptr;
        const uint32_t vb = 1u << (w & 31u);
        if (vw ? !(gcq::ld_cached(vw) & vb) : rd(dp) == -1) {
          int32_t expect = -1;
          disc = gcq::cas_rlx(dp, expect, odn);
          if (disc && vw) gcq::or_rlx(vw, vb);
        }
      }
      const uint64_t dm = __ballot(disc);
      // One activity store per run of discovering lanes with the same query
      // (neighbours of one vertex sit on consecutive lanes): a lane stores
      // only if the nearest discovering lane below it belongs to another query.
      const uint64_t below = dm & ((1ull << lane) - 1ull);
      const int pl = below ? 63 - __builtin_clzll(below) : (int)lane;
      const uint32_t pslot = __shfl(oslot, pl, 64);
      if (disc && (!below || pslot != oslot)) gcq::st_rlx(&d.alive[oslot * kAliveStride], (uint32_t)(L + 1));
      if (disc) stage[sc + gcq::rank_in(dm)] = (oslot << d.vbits) | w;
      sc += (uint32_t)__builtin_popcountll(dm);
      disc_n += disc ? 1u : 0u;
      while (sc >= d.flush_at) flush(qout, tid, stage, sc, sc < 64u ? sc : 64u, d.err);
    }
    if (one_round) break;  // every ticket of the level was handed out in this round
  }
  if (sc) flush(qout, tid, stage, sc, sc, d.err);
  return disc_n;
}

// The grid's scan of a level's posted vertices, per-query form (between the
// level's two barriers): each edge is probed and claimed as expand() does.
// Returns the calling lane's discoveries.
template <class F, bool kLive = true>
__device__ uint32_t hubs(const F& qout, const Dev& d, HubPool* hp, int32_t L, uint32_t tid, uint32_t nwaves,
                         uint32_t* stage) {
  auto rd = [](const auto* p) { return kLive ? gcq::ld_rlx(p) : gcq::ld_cached(p); };
  const uint32_t lane = lane_id();
  uint32_t sc = 0, disc_n = 0;
  hub_scan(hp, tid, nwaves, [&](const HubEntry* e, int32_t first, int32_t stop) {
    const uint32_t item = gcq::ld_cached(&e->item);
    const uint32_t slot = item >> d.vbits;
    const Slot* s = &d.slots[slot];
    const int32_t* adj = (const int32_t*)rd(&s->idx) + gcq::ld_cached(&e->beg);
    int32_t* const dep = (int32_t*)rd(&s->depth);
    uint32_t* const vis = kLive ? nullptr : (uint32_t*)rd(&s->visited);
    const int32_t dn = L - (int32_t)rd(&s->start) + 1;
    int32_t e0 = first + (int32_t)lane;
    uint32_t w_next = e0 < stop ? (uint32_t)adj[e0] : 0u;
    for (; e0 - (int32_t)lane < stop; e0 += 64) {
      const uint32_t w = w_next;
      if (e0 + 64 < stop) w_next = (uint32_t)adj[e0 + 64];
      bool disc = false;
      if (e0 < stop) {
        uint32_t* vw = vis ? vis + (w >> 5) : nullptr;
        const uint32_t vb = 1u << (w & 31u);
        if (vw ? !(gcq::ld_cached(vw) & vb) : rd(dep + w) == -1) {
          int32_t expect = -1;
          disc = gcq::cas_rlx(dep + w, expect, dn);
          if (disc && vw) gcq::or_rlx(vw, vb);
        }
      }
      const uint64_t dm = __ballot(disc);
      if (dm && lane == (uint32_t)__builtin_ctzll(dm)) gcq::st_rlx(&d.alive[slot * kAliveStride], (uint32_t)(L + 1));
      if (disc) stage[sc + gcq::rank_in(dm)] = (slot << d.vbits) | w;
      sc += (uint32_t)__builtin_popcountll(dm);
      disc_n += disc ? 1u : 0u;
      while (sc >= d.flush_at) flush(qout, tid, stage, sc, sc < 64u ? sc : 64u, d.err);
    }
  });
  if (sc) flush(qout, tid, stage, sc, sc, d.err);
  return disc_n;
}

}  // namespace bk

// ---------------------------------------------------------------- multi-source groups
namespace bk {

// Bit-parallel multi-source BFS (the MS-BFS of Then et al., "The More the
// Merrier: Efficient Multi-Source Graph Traversal", PVLDB 8(4), 2014) on the
// queue-driven, level-synchronous frontier.  Up to 64 queries on one graph
// form a group; a work item is (group, vertex), and the vertex's next-level
// mask holds the members that reach it at that level, so an edge is scanned
// once per level for the whole group instead of once per query.  Per vertex
// v of the group's graph:
//   seen[v]    members that have reached v, written only by the lane that
//              settles v (a vertex is queued at most once per level);
//   nxt[p][v]  members arriving at the next level of parity p, gathered by
//              fetch-or; the fetch-or that makes it non-zero queues (g, v).
// Settling v at level L: fresh = nxt[(L+1)&1][v] & ~seen[v] (the word is
// cleared for the level after next), each member in fresh gets depth
// L - start at v, and v's neighbours w receive fresh & ~seen[w] (seen[w] may
// be stale: a bit w already has is dropped again when w settles).  A bit
// moves one edge per level and only its first arrival at v is kept, so every
// member's depth array equals a plain BFS's bit for bit.
//
// kLive (the server): members join and leave while the kernel runs.  A
// member's bit is cleared from seen[] (atomic and, during the level after
// its admission) before it is seeded, so settling takes its fresh bits with
// one fetch-or on seen[v] (a load and a store could write a cleared bit
// back), and cached reads of seen[] are current from the next barrier on
// (every block drops its cached lines after a barrier).  Each level's fresh
// members are gathered per group (one fetch-or per wave and group) so that a
// member with nothing fresh at a level is known to be finished.
struct Group {
  uint64_t off, idx;  // CSR of the group's graph
  uint64_t seen;      // uint64_t[nv]
  uint64_t nxt[2];    // uint64_t[nv] each
  uint64_t depth;     // uint64_t[64]: each member's depth array (int32_t[nv])
  uint64_t start;     // int32_t[64]: level at which each member's source settles (0 = all members at 0)
  uint64_t fresh;     // kLive: uint64_t[2], the members settled at a level of each parity
};
static_assert(sizeof(Group) == 64, "group");

struct MsDev {
  const Group* groups;
  uint32_t* err;  // bit 0: a frontier enqueue failed
  uint32_t ngroups, vbits, pop_width, flush_at;
};

// The grid's scan of a level's posted vertices, multi-source form: each
// edge sends the posted vertex's fresh members to the target's gathering
// word.  Returns the calling lane's discoveries.
template <class F>
__device__ uint32_t ms_hubs(const F& qout, const MsDev& d, HubPool* hp, int32_t L, uint32_t tid, uint32_t nwaves,
                            uint32_t* stage) {
  const uint32_t lane = lane_id();
  uint32_t sc = 0, disc_n = 0;
  hub_scan(hp, tid, nwaves, [&](const HubEntry* e, int32_t first, int32_t stop) {
    const uint32_t item = gcq::ld_cached(&e->item);
    const int32_t beg = gcq::ld_cached(&e->beg);
    const uint64_t fresh = gcq::ld_cached(&e->fresh);
    const uint32_t g = item >> d.vbits;
    const Group* G = &d.groups[g];
    const int32_t* idx = (const int32_t*)gcq::ld_cached(&G->idx) + beg;
    const uint64_t* seen = (const uint64_t*)gcq::ld_cached(&G->seen);
    uint64_t* out = (uint64_t*)gcq::ld_cached(&G->nxt[L & 1]);
    int32_t e0 = first + (int32_t)lane;
    uint32_t w_next = e0 < stop ? (uint32_t)idx[e0] : 0u;
    for (; e0 - (int32_t)lane < stop; e0 += 64) {
      const uint32_t w = w_next;
      if (e0 + 64 < stop) w_next = (uint32_t)idx[e0 + 64];
      bool disc = false;
      if (e0 < stop) {
        const uint64_t send = fresh & ~gcq::ld_cached(seen + w);
        if (send) disc = gcq::or_rlx(out + w, send) == 0ull;
      }
      const uint64_t dm = __ballot(disc);
      if (disc) stage[sc + gcq::rank_in(dm)] = (g << d.vbits) | w;
      sc += (uint32_t)__builtin_popcountll(dm);
      disc_n += disc ? 1u : 0u;
      while (sc >= d.flush_at) flush(qout, tid, stage, sc, sc < 64u ? sc : 64u, d.err);
    }
  });
  if (sc) flush(qout, tid, stage, sc, sc, d.err);
  return disc_n;
}

// As expand(), for group items: counted pops, grid-position tickets on
// one-round levels, wave-cooperative scan of the settled vertices' edges.
template <class F, bool kLive = false>
__device__ uint32_t ms_expand(const F& qin, const F& qout, const MsDev& d, int32_t L, uint32_t tid, uint32_t* stage,
                              uint64_t h0, uint64_t grid_waves, uint64_t* end_out, HubPool* hp) {
  const uint32_t lane = lane_id();
  const uint32_t vmask = (1u << d.vbits) - 1u;
  uint32_t sc = 0, disc_n = 0;
  const bool pops = lane < d.pop_width;
  const uint64_t tix = h0 + (uint64_t)lane * grid_waves + tid / 64;
  uint64_t first_word = 0;
  if constexpr (F::kHeadFromDrain) {
    if (h0 != kNoHead && pops) first_word = qin.peek(tix);
  }
  const uint64_t end = qin.level_end();
  *end_out = end;
  const bool one_round = F::kHeadFromDrain && h0 != kNoHead && end >= h0 && end - h0 <= grid_waves * d.pop_width;
  for (;;) {
    uint32_t item = 0;
    int st = -1;
    if (pops) {
      if constexpr (F::kHeadFromDrain) {
        st = one_round ? qin.take(tid, tix, end, first_word, &item) : qin.pop_counted(tid, end, &item);
      } else {
        st = qin.pop_counted(tid, end, &item);
      }
    }
    if (__ballot(st >= 0) == 0) break;  // every popping lane drew a ticket past the level's end
    uint32_t g = 0;
    int32_t beg = 0, deg = 0;
    uint64_t idxp = 0, seenp = 0, outp = 0, fresh = 0;
    if (st > 0) {
      g = item >> d.vbits;
      const uint32_t v = item & vmask;
      const Group* G = &d.groups[g];  // fixed for the launch: read through the caches
      const int32_t* off = (const int32_t*)gcq::ld_cached(&G->off);
      idxp = gcq::ld_cached(&G->idx);
      seenp = gcq::ld_cached(&G->seen);
      outp = gcq::ld_cached(&G->nxt[L & 1]);
      uint64_t* const in = (uint64_t*)gcq::ld_cached(&G->nxt[(L + 1) & 1]) + v;
      uint64_t* const sp = (uint64_t*)seenp + v;
      const uint64_t arrived = gcq::ld_rlx(in);
      gcq::st_rlx(in, 0ull);  // this word gathers the level after next
      if constexpr (kLive) {
        fresh = arrived ? arrived & ~gcq::or_rlx(sp, arrived) : 0ull;
      } else {
        const uint64_t had = gcq::ld_rlx(sp);
        fresh = arrived & ~had;
        if (fresh) gcq::st_rlx(sp, had | fresh);
      }
      if (fresh) {
        const uint64_t* dtab = (const uint64_t*)gcq::ld_cached(&G->depth);
        const int32_t* stab = (const int32_t*)gcq::ld_cached(&G->start);
        for (uint64_t bits = fresh; bits; bits &= bits - 1) {
          const int b = __builtin_ctzll(bits);
          ((int32_t*)gcq::ld_cached(&dtab[b]))[v] = stab ? L - gcq::ld_cached(&stab[b]) : L;
        }
        beg = off[v];
        deg = off[v + 1] - beg;
        if (deg > kHubDeg && hub_post(hp, item, beg, deg, fresh)) deg = 0;  // the grid scans it
      }
    }
    if constexpr (kLive) {
      // this level's fresh members, one fetch-or per group present in the wave
      uint64_t pend = __ballot(fresh != 0ull);
      while (pend) {
        const int l0 = __builtin_ctzll(pend);
        const uint32_t gs = __shfl(g, l0, 64);
        const bool mine = fresh != 0ull && g == gs;
        uint64_t m = mine ? fresh : 0ull;
        for (int o = 32; o > 0; o >>= 1) m |= shfl64(m, (int)(lane ^ (uint32_t)o));
        if ((int)lane == l0) gcq::or_rlx((uint64_t*)gcq::ld_cached(&d.groups[gs].fresh) + (L & 1), m);
        pend &= ~__ballot(mine);
      }
    }
    int32_t incl = deg;  // inclusive prefix of degrees across the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += t;
    }
    const int32_t excl = incl - deg;
    const int32_t total = __shfl(incl, 63, 64);
    auto owner = [&](int32_t base) {  // first lane whose inclusive prefix exceeds this lane's edge
      const int32_t e = base + (int32_t)lane;
      const int32_t ec = e < total ? e : total - 1;
      int lo = 0, hi = 63;
      for (int s = 0; s < 6; ++s) {
        const int mid = (lo + hi) >> 1;
        if (__shfl(incl, mid, 64) > ec) hi = mid;
        else lo = mid + 1;
      }
      return lo;
    };
    auto neighbour = [&](int32_t base, int lo) -> uint32_t {  // this lane's edge target, 0 past the end
      const int32_t e = base + (int32_t)lane;
      const int32_t obeg = __shfl(beg, lo, 64);
      const int32_t oexcl = __shfl(excl, lo, 64);
      const uint64_t oidx = shfl64(idxp, lo);
      return e < total ? (uint32_t)((const int32_t*)oidx)[obeg + (e - oexcl)] : 0u;
    };
    int lo_next = total > 0 ? owner(0) : 0;
    uint32_t w_next = total > 0 ? neighbour(0, lo_next) : 0u;
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t e = base + (int32_t)lane;
      const int lo = lo_next;
      const uint32_t w_here = w_next;
      if (base + 64 < total) {
        lo_next = owner(base + 64);
        w_next = neighbour(base + 64, lo_next);
      }
      const uint32_t og = __shfl(g, lo, 64);
      const uint64_t ofresh = shfl64(fresh, lo);
      const uint64_t oseen = shfl64(seenp, lo);
      const uint64_t oout = shfl64(outp, lo);
      bool disc = false;
      uint32_t w = 0;
      if (e < total) {
        w = w_here;
        const uint64_t send = ofresh & ~gcq::ld_cached((const uint64_t*)oseen + w);
        if (send) disc = gcq::or_rlx((uint64_t*)oout + w, send) == 0ull;
      }
      const uint64_t dm = __ballot(disc);
      if (disc) stage[sc + gcq::rank_in(dm)] = (og << d.vbits) | w;
      sc += (uint32_t)__builtin_popcountll(dm);
      disc_n += disc ? 1u : 0u;
      while (sc >= d.flush_at) flush(qout, tid, stage, sc, sc < 64u ? sc : 64u, d.err);
    }
    if (one_round) break;  // every ticket of the level was handed out in this round
  }
  if (sc) flush(qout, tid, stage, sc, sc, d.err);
  return disc_n;
}

}  // namespace bk
