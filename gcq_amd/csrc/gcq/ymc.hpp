// ymc.hpp — WFQ: Yang & Mellor-Crummey wait-free queue ("A Wait-free Queue
// as Fast as Fetch-and-Add", PPoPP'16) over a flat HBM cell array.
//
// Algorithm follows the reference wfqueue_hip_opt.hpp:
//   sentinels ⊥=0, ⊤=~0, EMPTY=0, RETRY=~1           wfqueue_hip_opt.hpp:42-58
//   request state = pending<<63 | id                   :125-129
//   enq_fast / enq_slow / enq_commit / help_enq        :259-413
//   deq_fast / help_deq / deq_slow                     :416-524
//   wf_enqueue / wf_dequeue                            :527-555
// Own MI355X design choices:
//  * cells live in one flat array indexed by cell id (the reference's
//    segment pool is flat too, :208-237, but pays an atomicMax on every
//    access to track the high-water mark; here the mark is max(T, H)).
//  * a cell is 16 bytes: value (u64) and the enqueue / dequeue request
//    references as 32-bit words (0 = none, 1 = ⊤, k >= 2 = thread k-2).
//  * the pool is sized from a cell budget in HBM (288 GB per MI355X); an
//    exhausted pool makes enqueue return false and dequeue EMPTY, as in the
//    reference, and the host resets only the touched prefix.
//  * T and H tickets of the fast paths come from one aggregated FAA per wave.
#pragma once
#include "prim.hpp"

namespace gcq {
namespace ymc {

constexpr uint64_t kBottom = 0ull;                 // ⊥ value
constexpr uint64_t kTop = 0xFFFFFFFFFFFFFFFFull;   // ⊤ value
constexpr uint64_t kEmpty = 0ull;
constexpr uint64_t kRetry = 0xFFFFFFFFFFFFFFFEull;
constexpr uint32_t kRefNull = 0u, kRefTop = 1u;    // request references
constexpr uint32_t kHdrBytes = 2048;

struct Cell {
  uint64_t val;
  uint32_t enq;  // request reference
  uint32_t deq;
};
static_assert(sizeof(Cell) == 16, "cell");

// Per-thread handle (reference wf_handle, :161-178).  Requests are read by
// helpers; each handle is one 64-B half line.
struct alignas(64) Handle {
  uint64_t enq_val;     // published before enq_state
  uint64_t enq_state;   // pending<<63 | id
  uint64_t deq_id;
  uint64_t deq_state;
  uint64_t enq_id_last;
  uint32_t enq_peer, deq_peer;
  uint32_t next;
  uint32_t pad[3];
};
static_assert(sizeof(Handle) == 64, "handle");

struct Geometry {
  uint64_t cells;       // pool size
  uint32_t nthreads;    // handles
  uint32_t patience;
  uint32_t off_t, off_h, magic, pad;
  uint64_t ext_cells;   // 0: cells follow the header; else their own allocation
  uint64_t ext_handles; // handle array, for callers that pass only their own handle
};
constexpr uint32_t kMagic = 0x594D4331u;  // "YMC1"
struct Queue {
  Geometry g;
  uint8_t pad_[kHdrBytes - sizeof(Geometry)];
};
static_assert(sizeof(Queue) == kHdrBytes, "header");

__host__ __device__ constexpr uint64_t queue_bytes(uint64_t cells) { return kHdrBytes + 16ull * cells; }
__device__ __forceinline__ Cell* cells(Queue* q) {
  const uint64_t ext = q->g.ext_cells;
  return ext ? (Cell*)ext : (Cell*)((char*)q + kHdrBytes);
}
__device__ __forceinline__ uint64_t* Tp(Queue* q) { return (uint64_t*)((char*)q + q->g.off_t); }
__device__ __forceinline__ uint64_t* Hp(Queue* q) { return (uint64_t*)((char*)q + q->g.off_h); }

__host__ __device__ __forceinline__ uint64_t pack_state(uint32_t pending, uint64_t id) {
  return ((uint64_t)pending << 63) | (id & ((1ull << 63) - 1));
}
__host__ __device__ __forceinline__ bool st_pending(uint64_t s) { return (s >> 63) != 0; }
__host__ __device__ __forceinline__ uint64_t st_id(uint64_t s) { return s & ((1ull << 63) - 1); }

struct NoProbe {
  __device__ __forceinline__ void step() {}
  __device__ __forceinline__ void after_ticket() {}
  __device__ __forceinline__ void slow_entry() {}
};

// Header + handles; the cell array must already be zero (⊥ / none).
__device__ inline void init_slice(Queue* q, Handle* h, uint64_t ncells, uint32_t nthreads, uint32_t patience,
                                  uint32_t gid, uint32_t stride) {
  for (uint32_t i = gid; i < nthreads; i += stride) {
    Handle x;
    x.enq_val = 0;
    x.enq_state = pack_state(0, 0);
    x.deq_id = 0;
    x.deq_state = pack_state(0, 0);
    x.enq_id_last = 0;
    x.next = (i + 1) % nthreads;
    x.enq_peer = x.next;
    x.deq_peer = x.next;
    x.pad[0] = x.pad[1] = x.pad[2] = 0;
    h[i] = x;
  }
  if (gid == 0) {
    Geometry g;
    g.cells = ncells;
    g.nthreads = nthreads;
    g.patience = patience;
    g.off_t = 1 * kLine;
    g.off_h = 5 * kLine;
    g.magic = kMagic;
    g.pad = 0;
    g.ext_cells = 0;
    g.ext_handles = 0;
    q->g = g;
    *(uint64_t*)((char*)q + g.off_t) = 0;
    *(uint64_t*)((char*)q + g.off_h) = 0;
  }
}

template <class Probe = NoProbe>
struct Ops {
  Queue* q;
  Handle* hs;  // all handles
  Handle* h;   // this thread's handle
  uint32_t self;
  uint64_t ncells;
  Probe* pr;

  __device__ __forceinline__ Cell* cell(uint64_t i) const { return i < ncells ? &cells(q)[i] : nullptr; }

  __device__ __forceinline__ uint32_t ref_of(uint32_t tid) const { return tid + 2u; }
  __device__ __forceinline__ Handle* handle_of(uint32_t ref) const { return &hs[ref - 2u]; }

  // Raise an end counter to at least cid (:192-199).  The reference loops a
  // CAS; T and H are plain counters, so one fetch_max does it with a single
  // request to the hot line (help_deq calls this for every cell it scans).
  __device__ __forceinline__ void advance_end(uint64_t* e, uint64_t cid) {
    if (ld_rlx(e) < cid) max_rlx(e, cid);  // usually already past: a load, no RMW
  }
  __device__ __forceinline__ void enq_commit(Cell* c, uint64_t v, uint64_t cid) {  // :259-267
    advance_end(Tp(q), cid + 1);
    xchg_rlx(&c->val, v);
  }
  __device__ __forceinline__ bool try_claim(uint64_t* st, uint64_t id, uint64_t cid) {
    return cas_once(st, pack_state(1, id), pack_state(0, cid));
  }

  // ct: draw the T ticket together with the workgroup's other enqueuing
  // waves (group_faa).  Measured against per-wave FAAs at 65,536 threads
  // (32,768 for 75/25): split 25/75 level, 50/50 +9%, 75/25 +34%, balanced
  // -5%.  Combining idle consumers' claims as G-WFQ does gained WFQ
  // nothing (-6% at 16,384 threads), so its claims stay per wave.
  __device__ __forceinline__ bool enq_fast(uint64_t v, uint64_t* cid, Combine* ct = nullptr) {  // :269-281
    const uint64_t i = ct ? group_faa(ct, Tp(q), (uint64_t)1) : wave_faa(Tp(q), (uint64_t)1);
    pr->step();
    pr->after_ticket();
    Cell* c = cell(i);
    *cid = i;
    if (!c) return false;
    return cas_once(&c->val, kBottom, v);
  }

  __device__ bool enq_slow(uint64_t v, uint64_t cell_id) {  // :283-318
    pr->slow_entry();
    st_rlx(&h->enq_val, v);
    vm_wait();  // value acknowledged before the request is announced
    st_rlx(&h->enq_state, pack_state(1, cell_id));
    for (;;) {
      pr->step();
      const uint64_t i = faa_rlx(Tp(q), (uint64_t)1);
      Cell* c = cell(i);
      if (!c) break;
      if (cas_once(&c->enq, kRefNull, ref_of(self))) {
        const uint64_t cur = ld_rlx(&c->val);
        if (cur == kBottom || cur == kTop) {
          (void)try_claim(&h->enq_state, cell_id, i);
          break;
        }
      }
      if (!st_pending(ld_rlx(&h->enq_state))) break;
    }
    const uint64_t fid = st_id(ld_rlx(&h->enq_state));
    Cell* d = cell(fid);
    if (!d) return false;  // pool exhausted
    if (ld_rlx(&d->val) == kTop) enq_commit(d, v, fid);
    return true;
  }

  // Resolve cell i for a dequeuer: its value, EMPTY or RETRY (:320-413).
  __device__ uint64_t help_enq(Cell* c, uint64_t i) {
    uint64_t old = kBottom;
    if (!cas_rlx(&c->val, old, kTop) && old != kTop) return old;
    uint32_t ce = ld_rlx(&c->enq);
    if (ce == kRefNull) {
      uint32_t p = h->enq_peer;
      Handle* ph = &hs[p];
      uint64_t s = ld_rlx(&ph->enq_state);
      if (h->enq_id_last != 0 && h->enq_id_last != st_id(s)) {
        h->enq_id_last = 0;
        p = ph->next;
        h->enq_peer = p;
        ph = &hs[p];
        s = ld_rlx(&ph->enq_state);
      }
      if (st_pending(s) && st_id(s) > 0 && st_id(s) <= i) {
        uint32_t prev = kRefNull;
        cas_rlx(&c->enq, prev, ref_of(p));
        if (prev != kRefNull && prev != ref_of(p)) {
          ce = prev;
          h->enq_id_last = st_id(s);  // keep tracking this peer
        } else {
          ce = ref_of(p);
          h->enq_id_last = 0;
          h->enq_peer = ph->next;
        }
      } else {
        h->enq_id_last = 0;
        h->enq_peer = ph->next;
      }
      if (ce == kRefNull) {
        uint32_t prev = kRefNull;
        ce = cas_rlx(&c->enq, prev, kRefTop) ? kRefTop : prev;
      }
    }
    if (ce == kRefTop) return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    Handle* rh = handle_of(ce);
    const uint64_t s = ld_rlx(&rh->enq_state);
    vm_wait();  // the request state is read before its value
    const uint64_t v = ld_rlx(&rh->enq_val);
    if (!st_pending(s)) {
      if (st_id(s) == i) {
        if (ld_rlx(&c->val) == kTop) enq_commit(c, v, i);
        return ld_rlx(&c->val);
      }
      return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    }
    if (st_id(s) > i) return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    if (try_claim(&rh->enq_state, st_id(s), i) || ld_rlx(&rh->enq_state) == pack_state(0, i)) {
      if (ld_rlx(&c->val) == kTop) enq_commit(c, v, i);
    }
    return ld_rlx(&c->val);
  }

  // Head ticket bounded by the cells enqueuers have taken.  The wave leader
  // reads H, then T, and takes k = min(active lanes, T - H) tickets with one
  // FAA; lanes ranked below k get base + rank, the others claim again.  T <=
  // H at the leader's T read is EMPTY for every lane still claiming (the
  // early-empty rule of dequeue()).  Without the bound, a wave of 64 idle
  // consumers takes 64 cells for one pending item and marks 63 of them TOP,
  // and every enqueuer that later draws one of those cells has to retry.
  __device__ __forceinline__ bool claim_h(uint64_t* ticket) {
    for (;;) {
      const uint64_t m = active_mask();
      const uint32_t leader = first_lane(m);
      const uint32_t n = (uint32_t)__builtin_popcountll(m);
      uint64_t base = 0;
      uint32_t k = 0;
      if (lane_id() == leader) {
        const uint64_t h0 = ld_rlx(Hp(q));
        vm_wait();  // H is read before T
        const uint64_t t0 = ld_rlx(Tp(q));
        const uint64_t avail = t0 > h0 ? t0 - h0 : 0;
        k = avail < n ? (uint32_t)avail : n;
        if (k) base = faa_rlx(Hp(q), (uint64_t)k);
      }
      k = bcast32(k, leader);
      if (k == 0) return false;
      base = bcast64(base, leader);
      const uint32_t rank = rank_in(m);
      if (rank < k) {
        *ticket = base + rank;
        return true;
      }
    }
  }

  __device__ uint64_t deq_fast(uint64_t i) {  // :416-434, ticket from claim_h
    pr->step();
    pr->after_ticket();
    Cell* c = cell(i);
    if (!c) return kEmpty;
    // The ticket lies below T: its enqueuer is in flight.  Give it a moment
    // before help_enq marks the cell TOP and sends the enqueuer round again.
    for (uint32_t w = 0; w < kSlotWaitSpins && ld_rlx(&c->val) == kBottom; ++w) {
      __builtin_amdgcn_s_sleep(1);
      pr->step();
    }
    const uint64_t v = help_enq(c, i);
    if (v == kEmpty) return kEmpty;
    if (v != kTop && v != kRetry) {
      if (cas_once(&c->deq, kRefNull, kRefTop)) return v;
    }
    return kRetry;
  }

  __device__ void help_deq(Handle* helpee, uint32_t helpee_tid) {  // :436-501
    uint64_t s = ld_rlx(&helpee->deq_state);
    const uint64_t id = ld_rlx(&helpee->deq_id);
    if (!st_pending(s) || st_id(s) < id) return;
    uint64_t prior = id, i = id, cand = 0;
    for (;;) {
      pr->step();
      s = ld_rlx(&helpee->deq_state);
      if (!st_pending(s) || ld_rlx(&helpee->deq_id) != id) return;
      while (!cand && st_pending(s) && st_id(s) == prior) {
        Cell* c = cell(++i);
        if (!c) break;
        advance_end(Hp(q), i + 1);
        const uint64_t v = help_enq(c, i);
        if (v == kEmpty || (v != kTop && v != kRetry && ld_rlx(&c->deq) == kRefNull)) {
          cand = i;
          break;
        }
        s = ld_rlx(&helpee->deq_state);
        if (!st_pending(s) || ld_rlx(&helpee->deq_id) != id) return;
      }
      if (cand) {
        cas_once(&helpee->deq_state, pack_state(1, prior), pack_state(1, cand));
        cand = 0;
      }
      s = ld_rlx(&helpee->deq_state);
      if (!st_pending(s) || ld_rlx(&helpee->deq_id) != id) return;
      const uint64_t aidx = st_id(s);
      Cell* c = cell(aidx);
      if (!c) return;
      const uint32_t cur = ld_rlx(&c->deq);
      const uint32_t mine = ref_of(helpee_tid);
      if (ld_rlx(&c->val) == kTop || cas_once(&c->deq, kRefNull, mine) || cur == mine) {
        cas_once(&helpee->deq_state, s, pack_state(0, aidx));
        return;
      }
      prior = aidx;
      if (aidx >= i) i = aidx;
    }
  }

  __device__ uint64_t deq_slow(uint64_t cid) {  // :503-524
    pr->slow_entry();
    st_rlx(&h->deq_id, cid);
    vm_wait();
    st_rlx(&h->deq_state, pack_state(1, cid));
    help_deq(h, self);
    const uint64_t idx = st_id(ld_rlx(&h->deq_state));
    Cell* c = cell(idx);
    advance_end(Hp(q), idx + 1);
    if (!c) return kEmpty;
    const uint64_t v = ld_rlx(&c->val);
    return v == kTop ? kEmpty : v;
  }

  __device__ bool enqueue(uint64_t v, Combine* ct = nullptr) {  // :527-536
    for (int p = (int)q->g.patience; p >= 0; --p) {
      uint64_t cid = 0;
      if (enq_fast(v, &cid, ct)) return true;
      if (p == 0) return enq_slow(v, cid);
    }
    return false;
  }

  __device__ uint64_t dequeue() {  // :538-555
    // Every ticket comes from claim_h, which takes no cell when T <= H: then
    // every enqueue cell is already owned by a dequeuer (EMPTY at the T
    // read).  Idle consumers would otherwise burn a cell per attempt.
    uint64_t v = kRetry, cid = 0;
    for (int p = (int)q->g.patience; p >= 0; --p) {
      if (!claim_h(&cid)) return kEmpty;
      v = deq_fast(cid);
      if (v != kRetry) break;
    }
    if (v == kRetry) v = deq_slow(cid);
    if (v != kEmpty) {
      const uint32_t p = h->deq_peer;
      help_deq(&hs[p], p);
      h->deq_peer = hs[p].next;
    }
    return v;
  }
};

// `hs` is the handle array, `tid` the caller's thread id (< nthreads).
template <class Probe = NoProbe>
__device__ __forceinline__ bool enqueue(Queue* q, Handle* hs, uint32_t tid, uint64_t v, Probe* pr = nullptr,
                                        Combine* ct = nullptr) {
  NoProbe np;
  Ops<Probe> o{q, hs, &hs[tid], tid, q->g.cells, pr ? pr : (Probe*)&np};
  return o.enqueue(v, ct);
}
template <class Probe = NoProbe>
__device__ __forceinline__ uint64_t dequeue(Queue* q, Handle* hs, uint32_t tid, Probe* pr = nullptr) {
  NoProbe np;
  Ops<Probe> o{q, hs, &hs[tid], tid, q->g.cells, pr ? pr : (Probe*)&np};
  return o.dequeue();
}
// High-water mark of touched cells (for resets and pool-exhaustion checks).
__device__ __forceinline__ uint64_t high_water(Queue* q) {
  const uint64_t t = ld_rlx(Tp(q)), h = ld_rlx(Hp(q));
  return t > h ? t : h;
}

// Recycling a quiescent queue (no operation in flight and none starting
// until rewind() has completed): every thread of a grid calls
// recycle_cells to clear its share of the touched cells and reset_handle on
// its own handle (the owner reads its private fields with plain loads),
// then one thread calls rewind() once all of them are done.  Shared words
// are cleared with device-scope stores so waves on other XCDs see them.
__device__ inline void recycle_cells(Queue* q, uint64_t gid, uint64_t stride) {
  const uint64_t hw = min(high_water(q), q->g.cells);
  uint64_t* w = (uint64_t*)cells(q);
  for (uint64_t i = gid; i < 2 * hw; i += stride) st_rlx(&w[i], 0ull);
}
__device__ inline void reset_handle(Queue* q, Handle* hs, uint32_t tid) {
  const uint32_t n = q->g.nthreads;
  if (tid >= n) return;
  Handle* h = &hs[tid];
  st_rlx(&h->enq_val, 0ull);
  st_rlx(&h->enq_state, 0ull);
  st_rlx(&h->deq_id, 0ull);
  st_rlx(&h->deq_state, 0ull);
  h->enq_id_last = 0;
  h->enq_peer = h->deq_peer = (tid + 1) % n;
}
__device__ inline void rewind(Queue* q) {
  st_rlx(Tp(q), 0ull);
  st_rlx(Hp(q), 0ull);
}

}  // namespace ymc
}  // namespace gcq
