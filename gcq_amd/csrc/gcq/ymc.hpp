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
//  * Ring mode (library queues): the pool is a ring of P segments of S cells
//    and cell id i lives at i mod (P * S), so a bounded pool serves an
//    unbounded run.  The reference declares hazard pointers and a segment
//    floor for this (wf_handle::hzdp, wf_queue::I, wfqueue_hip_opt.hpp:
//    161-178) and never uses them; its pool runs dry (:217-220).  Here:
//     - each wave publishes a floor `lo`: every cell it may touch in its
//       current operation has id >= lo.  It is set to cand * S, where cand
//       is the last reclaim round's min(H, T) / S, read at the operation's
//       start (tickets only grow, so every later ticket of either counter is
//       above it).  The first publication of a wave in a kernel completes
//       (s_waitcnt) before its first ticket; later ones only raise it.  A
//       helper about to scan an older request's cells lowers it to that
//       request's id, waits for the store, and then backs off if a reclaim
//       has already promised those segments away (cand above them);
//     - a reclaim round (one wave at a time, under a lock) reads H then T,
//       publishes cand = min(H, T) / S, waits, scans the waves' floors, and
//       frees the segments below min(cand, floors): it clears their cells
//       and only then raises `avail`, the first logical segment not yet
//       usable (the freed physical segments now carry the next lap);
//     - a ticket in a segment at or above avail waits for a round to free
//       room (boundedly: past kWindowSpins checks the operation fails, as an
//       exhausted pool does in the reference), and a wave whose ticket
//       reaches `next_try` starts a round.
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

// A cell is 16 B in two arrays indexed alike: its value (u64) in the first,
// its enqueue and dequeue request references (u32 each) in the second.  An
// enqueue's fast path touches only the value and a dequeue's the value and
// the dequeue reference, so the 64 consecutive cells of a wave's tickets
// span 4 lines of each array (a device-scope atomic is one request per line
// per wave instruction) instead of 8 lines of interleaved 16-B cells.
struct Refs {
  uint32_t enq, deq;
};
static_assert(sizeof(Refs) == 8, "refs");
struct Cell {
  uint64_t* val;
  uint32_t* enq;
  uint32_t* deq;
  __device__ explicit operator bool() const { return val != nullptr; }
};

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
  uint32_t ring;        // 1: the pool is a ring of segments, reclaimed as the run goes
  uint32_t seg_log2;    // ring: log2 of the cells per segment
  uint32_t off_win, off_bound, off_lock, pad2;  // off_bound: the clearing job
};

// Ring mode's window line, read at the start of every operation (it changes
// once or twice per reclaim round).  All values are logical segment numbers.
struct Window {
  uint64_t avail;     // segments below this are usable
  uint64_t next_try;  // a ticket in this segment or above starts a reclaim round
  uint64_t cand;      // the last round's min(H, T) / S
  uint64_t pad[13];
};
static_assert(sizeof(Window) == kLine, "one line");
constexpr uint64_t kIdle = ~0ull;            // a wave floor that holds nothing
constexpr uint64_t kWithdrawn = (1ull << 63) - 1;  // enqueue request given up (pool exhausted)
constexpr uint32_t kWindowSpins = 1u << 16;  // checks of `avail` before a ticket gives up
constexpr uint32_t kPatientRounds = 16;
constexpr uint32_t kChainRounds = 4;
constexpr uint32_t kClearPerLane = 32;     // cells per active lane in a clearing chunk (2,048 for a full wave)
constexpr uint32_t kClearChunkLog2 = 11;  // a full wave's chunk, for sizing a round's share
// Segments one reclaim round may free (its clearing job): a quarter of the
// ring, at most GCQ_YMC_ROUND_SEGS.
#ifndef GCQ_YMC_ROUND_SEGS
#define GCQ_YMC_ROUND_SEGS 512
#endif
__host__ __device__ constexpr uint64_t kRoundSegs(uint64_t P) {
  return P / 4 > GCQ_YMC_ROUND_SEGS ? GCQ_YMC_ROUND_SEGS : (P / 4 ? P / 4 : 1);
}
constexpr uint32_t kMagic = 0x594D4331u;  // "YMC1"
struct Queue {
  Geometry g;
  uint8_t pad_[kHdrBytes - sizeof(Geometry)];
};
static_assert(sizeof(Queue) == kHdrBytes, "header");

__host__ __device__ constexpr uint64_t queue_bytes(uint64_t cells) { return kHdrBytes + 16ull * cells; }
// The values: g.cells u64 words after the header (or at ext_cells), then
// the reference pairs, g.cells more u64 words.
__device__ __forceinline__ uint64_t* vals(Queue* q) {
  const uint64_t ext = ld_const(&q->g).ext_cells;
  return ext ? (uint64_t*)ext : (uint64_t*)((char*)q + kHdrBytes);
}
__device__ __forceinline__ Refs* refs(Queue* q) { return (Refs*)(vals(q) + ld_const(&q->g).cells); }
__device__ __forceinline__ Cell cell_at(Queue* q, uint64_t k) {
  Refs* r = refs(q) + k;
  return Cell{vals(q) + k, &r->enq, &r->deq};
}
__device__ __forceinline__ uint64_t* Tp(Queue* q) { return (uint64_t*)((char*)q + ld_const(&q->g).off_t); }
__device__ __forceinline__ uint64_t* Hp(Queue* q) { return (uint64_t*)((char*)q + ld_const(&q->g).off_h); }
__device__ __forceinline__ Window* win(Queue* q) { return (Window*)((char*)q + ld_const(&q->g).off_win); }
// Ring mode's clearing job: the dead cells [0, end) (logical ids) are
// cleared in chunks claimed at `next`; `cleared` counts the cleared cells.
// A new job (a higher end) is published only once cleared == end.
struct ClearJob {
  uint64_t end, next, cleared, pad[13];
};
static_assert(sizeof(ClearJob) == kLine, "one line");
__device__ __forceinline__ ClearJob* job(Queue* q) { return (ClearJob*)((char*)q + ld_const(&q->g).off_bound); }
__device__ __forceinline__ uint32_t* lockp(Queue* q) { return (uint32_t*)((char*)q + ld_const(&q->g).off_lock); }
// Ring mode: one published floor per wave (thread id / 64), after the handles.
__host__ __device__ constexpr uint64_t wave_slots(uint32_t nthreads) { return (nthreads + 63u) / 64u; }
__device__ __forceinline__ uint64_t* wave_lo(Queue* q, Handle* hs) { return (uint64_t*)(hs + ld_const(&q->g).nthreads); }
// Per wave of a workgroup, in LDS: whether the wave has published its floor
// in this kernel, and the value it published (0 after a helper lowered it).
// avail and next_try are cached here too and read again every
// kWindowRefresh operations (or when a ticket reaches the cached value): a
// stale one only makes a ticket look the line up again.  cand is read on
// every operation: a wave's floor must stay within reach of the counters,
// or a dequeuer waiting at the window's end would hold the window back
// itself.
constexpr uint32_t kWindowRefresh = 16;
struct WaveLds {
  uint64_t lo[16];
  uint64_t avail[16], next_try[16];
  uint64_t last_t[16];  // the wave's last T ticket: a lower bound of T
  uint8_t pub[16], age[16], low[16];
};
__device__ __forceinline__ WaveLds* wave_lds() {
  __shared__ WaveLds s_wl;
  return &s_wl;
}

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
    g.ring = 0;
    g.seg_log2 = 0;
    g.off_win = 9 * kLine;
    g.off_bound = 11 * kLine;
    g.off_lock = 13 * kLine;
    g.pad2 = 0;
    q->g = g;
    *(uint64_t*)((char*)q + g.off_t) = 0;
    *(uint64_t*)((char*)q + g.off_h) = 0;
  }
}

// Ring mode on top of init_slice (ncells a power of two, >= 4096; the cells
// zero; the handle allocation followed by wave_slots(nthreads) floors).
__host__ __device__ constexpr uint32_t ring_seg_log2(uint64_t ncells) {
  uint32_t l = 0;
  while ((1ull << (l + 1)) <= ncells / 64 && l < 10) ++l;  // S = min(1024, cells / 64): at least 64 segments
  return l;
}
__device__ inline void init_ring_slice(Queue* q, Handle* hs, uint64_t ncells, uint32_t nthreads, uint32_t gid,
                                       uint32_t stride) {
  uint64_t* lo = (uint64_t*)(hs + nthreads);
  for (uint64_t i = gid; i < wave_slots(nthreads); i += stride) lo[i] = kIdle;
  if (gid == 0) {
    const uint32_t sl = ring_seg_log2(ncells);
    q->g.ring = 1;
    q->g.seg_log2 = sl;
    Window* w = (Window*)((char*)q + q->g.off_win);
    const uint64_t P = ncells >> sl;
    w->avail = P;
    w->next_try = kRoundSegs(P) / 2;
    w->cand = 0;
    ClearJob* j = (ClearJob*)((char*)q + q->g.off_bound);
    j->end = 0;
    j->next = 0;
    j->cleared = 0;
    *(uint32_t*)((char*)q + q->g.off_lock) = 0;
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
  // ring mode, loaded by prologue()
  uint32_t ring = 0, seg_log2 = 0;
  uint64_t avail = 0, next_try = 0, cand = 0;

  // patient: a dequeuer's own ticket.  A dequeuer that gave its ticket up
  // could orphan the item its enqueuer writes into that cell later, so it
  // waits kPatientRounds times longer than an enqueuer (which may give its
  // ticket up: the cell's dequeuer then marks the cell).
  __device__ __forceinline__ Cell cell(uint64_t i, bool patient = false) {
    if (!ring) return i < ncells ? cell_at(q, i) : Cell{nullptr, nullptr, nullptr};
    if ((i >> seg_log2) >= avail && !wait_window(i >> seg_log2, patient ? kPatientRounds : 1u))
      return Cell{nullptr, nullptr, nullptr};
    return cell_at(q, i & (ncells - 1));
  }

  // ---------------------------------------------------------------- ring mode
  // Start of an operation, all active lanes: load the window line; a wave's
  // first operation in a kernel publishes its floor and waits for the store
  // to complete before it takes any ticket.
  __device__ __forceinline__ void prologue() {
    ring = ld_const(&q->g).ring;
    if (!ring) return;
    seg_log2 = ld_const(&q->g).seg_log2;
    const uint32_t L = first_lane(active_mask());
    uint64_t a = 0, nt = 0, cd = 0;
    if (lane_id() == L) {
      WaveLds* s = wave_lds();
      const uint32_t wv = wave_in_block();
      const bool first = !(s->pub[wv] & 1u);
      Window* w = win(q);
      cd = ld_rlx(&w->cand);
      if (first || s->age[wv] == 0) {
        a = ld_rlx(&w->avail);
        nt = ld_rlx(&w->next_try);
        s->avail[wv] = a;
        s->next_try[wv] = nt;
        s->age[wv] = (uint8_t)kWindowRefresh;
      } else {
        a = s->avail[wv];
        nt = s->next_try[wv];
        s->age[wv] = (uint8_t)(s->age[wv] - 1u);
      }
      if (first) {
        const uint64_t lo = cd << seg_log2;
        st_rlx(&wave_lo(q, hs)[self / kWave], lo);
        vm_wait();  // published before this wave's first ticket
        s->lo[wv] = lo;
        s->low[wv] = 0;
        s->last_t[wv] = lo;
        s->pub[wv] = 1;
      }
    }
    avail = bcast64(a, L);
    next_try = bcast64(nt, L);
    // cand stays in the leader lane, which alone uses it (at the epilogue):
    // no broadcast, so the operation does not wait for the load.
    cand = cd;
  }
  // End of an operation, all active lanes: the wave's floor for its next
  // operation is the cand read at this one's start (every later ticket of
  // either counter is at least cand * S).
  __device__ __forceinline__ void epilogue() {
    if (!ring) return;
    const uint32_t L = first_lane(active_mask());
    if (lane_id() == L) {
      WaveLds* s = wave_lds();
      const uint32_t wv = wave_in_block();
      const uint64_t lo = cand << seg_log2;
      if (s->lo[wv] != lo || s->low[wv]) {
        st_rlx(&wave_lo(q, hs)[self / kWave], lo);
        s->lo[wv] = lo;
        s->low[wv] = 0;
      }
    }
  }
  // A helper about to touch the cells of a request that starts at cell x:
  // lower the wave's floor to x and wait for the store; false when a reclaim
  // round has already promised x's segment away (cand above it), so the
  // helper must leave that request alone.
  __device__ __forceinline__ bool protect(uint64_t x) {
    if (!ring) return true;
    WaveLds* s = wave_lds();
    const uint32_t wv = wave_in_block();
    // x at or above the wave's published floor, or (in the lane that read
    // it at this operation's start) within reach of cand, is covered
    if ((x >> seg_log2) >= cand || x >= __hip_atomic_load(&s->lo[wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      return true;
    __hip_atomic_fetch_min(&wave_lo(q, hs)[self / kWave], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_min(&s->lo[wv], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    s->low[wv] = 1;  // the epilogue raises the floor again
    vm_wait();
    return (x >> seg_log2) >= ld_rlx(&win(q)->cand);
  }
  // Wait (boundedly) until segment seg is usable; lanes may be divergent.
  __device__ bool wait_window(uint64_t seg, uint32_t rounds = 1) {
    for (uint64_t k = 0; k < (uint64_t)kWindowSpins * rounds; ++k) {
      const uint64_t a = ld_rlx(&win(q)->avail);
      if (seg < a) {
        avail = a;
        wave_lds()->avail[wave_in_block()] = a;  // any lane: every value read is a valid lower bound
        return true;
      }
      if ((k & 63u) == 0 || help_clear(1)) reclaim();
      __builtin_amdgcn_s_sleep(2);
    }
    return false;
  }
  // After a wave-level ticket: start a round when the wave's tickets reach
  // next_try (all active lanes).
  __device__ __forceinline__ void maybe_reclaim(uint64_t ticket) {
    if (!ring || !__ballot((ticket >> seg_log2) >= next_try)) return;
    const uint32_t L = first_lane(active_mask());
    uint64_t nt = 0;
    if (lane_id() == L) {  // the cached next_try may be stale
      nt = ld_rlx(&win(q)->next_try);
      wave_lds()->next_try[wave_in_block()] = nt;
    }
    next_try = bcast64(nt, L);
    if (__ballot((ticket >> seg_log2) >= next_try)) reclaim_and_clear(kRoundSegs(ncells >> seg_log2) << seg_log2 >> kClearChunkLog2);
  }
  // One reclaim round by the calling wave's active lanes, if the lock is free
  // and the last clearing job is finished: publish cand, read the floors,
  // and publish the dead range as the next clearing job (clr_end), then
  // help clear it.
  __device__ void reclaim() {
    const uint64_t m = active_mask();
    const uint32_t L = first_lane(m), n = (uint32_t)__builtin_popcountll(m), r = rank_in(m);
    uint32_t won = 0;
    if (lane_id() == L && ld_rlx(lockp(q)) == 0u) {
      uint32_t e = 0;
      won = cas_rlx(lockp(q), e, 1u) ? 1u : 0u;
    }
    if (!bcast32(won, L)) return;
    Window* w = win(q);
    ClearJob* j = job(q);
    const uint64_t P = ncells >> seg_log2;
    uint64_t b = 0, old_end = 0;
    uint32_t busy = 0;
    if (lane_id() == L) {
      const uint64_t hh = ld_rlx(Hp(q));
      vm_wait();  // H read before T
      const uint64_t tt = ld_rlx(Tp(q));
      b = (hh < tt ? hh : tt) >> seg_log2;
      if (b > ld_rlx(&w->cand)) st_rlx(&w->cand, b);
      vm_wait();  // cand published before the floors are read
      old_end = ld_rlx(&j->end);
      busy = ld_rlx(&j->cleared) != old_end;  // the last job is still being cleared
    }
    b = bcast64(b, L);
    old_end = bcast64(old_end, L) >> seg_log2;
    uint64_t nb = old_end;
    if (!bcast32(busy, L)) {
      const uint64_t* lo = wave_lo(q, hs);
      uint64_t mn = kIdle;
      for (uint64_t k = r; k < wave_slots(ld_const(&q->g).nthreads); k += n) {
        const uint64_t v = ld_rlx(&lo[k]);
        mn = v < mn ? v : mn;
      }
      for (uint64_t bits = m; bits; bits &= bits - 1) {
        const uint64_t v = bcast64(mn, first_lane(bits));
        mn = v < mn ? v : mn;
      }
      nb = mn == kIdle ? b : (mn >> seg_log2);
      if (nb > b) nb = b;
      if (nb > old_end + kRoundSegs(P)) nb = old_end + kRoundSegs(P);
      // The segments [old_end, nb) are dead: they become the clearing job.
      if (lane_id() == L && nb > old_end) st_rlx(&j->end, nb << seg_log2);
    }
    if (lane_id() == L) {
      // the next round once the tickets have moved half a round on
      const uint64_t step = kRoundSegs(P) / 2 ? kRoundSegs(P) / 2 : 1;
      const uint64_t t1 = nb + step, t2 = b + step;
      st_rlx(&w->next_try, t1 > t2 ? t1 : t2);
      vm_wait();
      st_rlx(lockp(q), 0u);
    }
  }
  // A reclaim round, then help clear its job; a wave that finishes a job
  // starts the next round at once (at most kChainRounds times).
  __device__ void reclaim_and_clear(uint64_t chunks) {
    for (uint32_t k = 0; k < kChainRounds; ++k) {
      reclaim();
      if (!help_clear(chunks)) return;  // no job finished: done for now
    }
  }

  // Clear up to `chunks` chunks of the current job (any active lanes).  A
  // chunk is claimed by a CAS on the job's cursor, cleared with device-scope
  // stores (like every other access to a cell: clearing with plain stores
  // and a release fence lost items, 58 in eight runs of 10^9 operations, 4
  // with an added invalidate, because a stale copy of a cleared line in
  // some XCD's L2 outlived the clear), and counted.  The chunk that brings
  // the count to the job's end completes the job: every dead cell is clear,
  // and the window moves to end / S + P.
  __device__ bool help_clear(uint64_t chunks) {
    const uint64_t m = active_mask();
    const uint32_t L = first_lane(m), n = (uint32_t)__builtin_popcountll(m), r = rank_in(m);
    ClearJob* j = job(q);
    uint64_t* const vw = vals(q);
    uint64_t* const rw = (uint64_t*)refs(q);
    for (uint64_t t = 0; t < chunks; ++t) {
      uint64_t c0 = 0, c1 = 0, e = 0;
      if (lane_id() == L) {
        e = ld_rlx(&j->end);  // fixed until this job is finished, so until this chunk is counted
        uint64_t c = ld_rlx(&j->next);
        // kClearPerLane cells per active lane: a wave with few lanes left
        // (a divergent waiter) takes a small chunk, so no chunk holds the
        // job back for long.
        const uint64_t chunk = (uint64_t)kClearPerLane * n;
        while (c < e) {
          const uint64_t nc = c + chunk < e ? c + chunk : e;
          if (cas_rlx(&j->next, c, nc)) {
            c0 = c;
            c1 = nc;
            break;
          }
        }
      }
      c0 = bcast64(c0, L);
      c1 = bcast64(c1, L);
      if (c1 == c0) return false;
      for (uint64_t c = c0 + r; c < c1; c += n) {
        st_rlx(&vw[c & (ncells - 1)], 0ull);
        st_rlx(&rw[c & (ncells - 1)], 0ull);
      }
      vm_wait();  // this chunk is clear before it is counted
      uint32_t fin = 0;
      if (lane_id() == L) {
        const uint64_t done = faa_rlx(&j->cleared, c1 - c0) + (c1 - c0);
        if (done == e) {
          max_rlx(&win(q)->avail, (e >> seg_log2) + (ncells >> seg_log2));
          fin = 1;
        }
      }
      if (bcast32(fin, L)) return true;
    }
    return false;
  }

  __device__ __forceinline__ uint32_t ref_of(uint32_t tid) const { return tid + 2u; }
  __device__ __forceinline__ Handle* handle_of(uint32_t ref) const { return &hs[ref - 2u]; }

  // Raise an end counter to at least cid (:192-199).  The reference loops a
  // CAS; T and H are plain counters, so one fetch_max does it with a single
  // request to the hot line (help_deq calls this for every cell it scans).
  __device__ __forceinline__ void advance_end(uint64_t* e, uint64_t cid) {
    if (ld_rlx(e) < cid) max_rlx(e, cid);  // usually already past: a load, no RMW
  }
  __device__ __forceinline__ void enq_commit(Cell c, uint64_t v, uint64_t cid) {  // :259-267
    advance_end(Tp(q), cid + 1);
    xchg_rlx(c.val, v);
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
    if (ring && lane_id() == first_lane(active_mask())) wave_lds()->last_t[wave_in_block()] = i;
    maybe_reclaim(i);
    Cell c = cell(i);
    *cid = i;
    if (!c) {
      *cid = kWithdrawn;  // no cell: the pool is exhausted
      return false;
    }
    return cas_once(c.val, kBottom, v);
  }

  __device__ bool enq_slow(uint64_t v, uint64_t cell_id) {  // :283-318
    pr->slow_entry();
    st_rlx(&h->enq_val, v);
    vm_wait();  // value acknowledged before the request is announced
    st_rlx(&h->enq_state, pack_state(1, cell_id));
    for (;;) {
      pr->step();
      const uint64_t i = faa_rlx(Tp(q), (uint64_t)1);
      Cell c = cell(i);
      if (!c) {
        // Pool exhausted: withdraw the request unless a helper has already
        // claimed a cell for it (then that cell receives the value below).
        if (cas_once(&h->enq_state, pack_state(1, cell_id), pack_state(0, kWithdrawn))) return false;
        break;
      }
      if (cas_once(c.enq, kRefNull, ref_of(self))) {
        const uint64_t cur = ld_rlx(c.val);
        if (cur == kBottom || cur == kTop) {
          (void)try_claim(&h->enq_state, cell_id, i);
          break;
        }
      }
      if (!st_pending(ld_rlx(&h->enq_state))) break;
    }
    const uint64_t fid = st_id(ld_rlx(&h->enq_state));
    Cell d = cell(fid);
    if (!d) return false;  // pool exhausted
    // The enqueuer writes its value into the cell claimed for its request
    // whatever the cell holds (⊥ or ⊤), as YMC's enq_slow does.  The
    // reference writes it only over ⊤ (wfqueue_hip_opt.hpp:308-314) and
    // otherwise leaves it to the cell's dequeuer, which reads the value from
    // the request; once this thread has announced its next slow-path request
    // that value is gone, and the item is lost.  No other enqueuer holds this
    // cell's ticket, and a helper commits the same value, so the store is
    // safe.
    enq_commit(d, v, fid);
    return true;
  }

  // Resolve cell i for a dequeuer: its value, EMPTY or RETRY (:320-413).
  __device__ uint64_t help_enq(Cell c, uint64_t i) {
    uint64_t old = kBottom;
    if (!cas_rlx(c.val, old, kTop) && old != kTop) return old;
    uint32_t ce = ld_rlx(c.enq);
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
        cas_rlx(c.enq, prev, ref_of(p));
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
        ce = cas_rlx(c.enq, prev, kRefTop) ? kRefTop : prev;
      }
    }
    if (ce == kRefTop) return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    Handle* rh = handle_of(ce);
    const uint64_t s = ld_rlx(&rh->enq_state);
    vm_wait();  // the request state is read before its value
    const uint64_t v = ld_rlx(&rh->enq_val);
    if (!st_pending(s)) {
      if (st_id(s) == i) {
        if (ld_rlx(c.val) == kTop) enq_commit(c, v, i);
        return ld_rlx(c.val);
      }
      return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    }
    if (st_id(s) > i) return ld_rlx(Tp(q)) <= i ? kEmpty : kRetry;
    if (try_claim(&rh->enq_state, st_id(s), i) || ld_rlx(&rh->enq_state) == pack_state(0, i)) {
      if (ld_rlx(c.val) == kTop) enq_commit(c, v, i);
    }
    return ld_rlx(c.val);
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
        const uint64_t room = t0 > h0 ? t0 - h0 : 0;
        k = room < n ? (uint32_t)room : n;
        if (k) base = faa_rlx(Hp(q), (uint64_t)k);
      }
      k = bcast32(k, leader);
      if (k == 0) return false;
      base = bcast64(base, leader);
      maybe_reclaim(base);
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
    Cell c = cell(i, true);
    if (!c) return kEmpty;
    // The ticket lies below T: its enqueuer is in flight.  Give it a moment
    // before help_enq marks the cell TOP and sends the enqueuer round again.
    for (uint32_t w = 0; w < kSlotWaitSpins && ld_rlx(c.val) == kBottom; ++w) {
      __builtin_amdgcn_s_sleep(1);
      pr->step();
    }
    const uint64_t v = help_enq(c, i);
    if (v == kEmpty) return kEmpty;
    if (v != kTop && v != kRetry) {
      if (cas_once(c.deq, kRefNull, kRefTop)) return v;
    }
    return kRetry;
  }

  __device__ void help_deq(Handle* helpee, uint32_t helpee_tid) {  // :436-501
    uint64_t s = ld_rlx(&helpee->deq_state);
    const uint64_t id = ld_rlx(&helpee->deq_id);
    if (!st_pending(s) || st_id(s) < id) return;
    if (helpee != h && !protect(id)) return;  // ring mode: the request's cells may be reclaimed
    uint64_t prior = id, i = id, cand = 0;
    for (;;) {
      pr->step();
      s = ld_rlx(&helpee->deq_state);
      if (!st_pending(s) || ld_rlx(&helpee->deq_id) != id) return;
      while (!cand && st_pending(s) && st_id(s) == prior) {
        Cell c = cell(++i);
        if (!c) break;
        advance_end(Hp(q), i + 1);
        const uint64_t v = help_enq(c, i);
        if (v == kEmpty || (v != kTop && v != kRetry && ld_rlx(c.deq) == kRefNull)) {
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
      Cell c = cell(aidx);
      if (!c) return;
      const uint32_t cur = ld_rlx(c.deq);
      const uint32_t mine = ref_of(helpee_tid);
      if (ld_rlx(c.val) == kTop || cas_once(c.deq, kRefNull, mine) || cur == mine) {
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
    Cell c = cell(idx);
    advance_end(Hp(q), idx + 1);
    if (!c) return kEmpty;
    const uint64_t v = ld_rlx(c.val);
    return v == kTop ? kEmpty : v;
  }

  // Ring mode, all active lanes, before an enqueue takes a ticket: false
  // when T is within kFullMargin segments of the window's end (the queue is
  // full).  A producer that took a ticket beyond the window would have to
  // wait for room there, and dequeuers that reach its cell would wait for
  // it, holding the window back themselves; refusing the enqueue instead
  // keeps every ticket inside the window (up to what other waves take
  // between the check and their FAA, which the margin covers).  The check
  // costs nothing until the wave's last T ticket comes near the end.
  __device__ bool room_to_enqueue() {
    if (!ring) return true;
    const uint64_t margin = (ld_const(&q->g).nthreads >> seg_log2) + 2;
    const uint32_t L = first_lane(active_mask());
    const uint64_t a0 = avail;  // wave-uniform
    uint64_t a = a0;
    uint32_t near = 0;
    if (lane_id() == L) {
      WaveLds* s = wave_lds();
      near = (s->last_t[wave_in_block()] >> seg_log2) + margin >= a0;
    }
    if (!bcast32(near, L)) return true;
    uint32_t ok = 0;
    for (int k = 0; k < 2; ++k) {
      if (lane_id() == L) {
        WaveLds* s = wave_lds();
        const uint32_t wv = wave_in_block();
        const uint64_t t = ld_rlx(Tp(q));
        a = ld_rlx(&win(q)->avail);
        s->avail[wv] = a;
        s->last_t[wv] = t;
        ok = (t >> seg_log2) + margin < a;
      }
      a = bcast64(a, L);
      ok = bcast32(ok, L);
      if (ok || k) break;
      // Try once to move the window: a round, and help clear its job (or the
      // unfinished one); no ticket is taken while the queue is full, so no
      // ticket-driven round would come.
      reclaim_and_clear(kRoundSegs(ncells >> seg_log2) << seg_log2 >> kClearChunkLog2);
    }
    avail = a;
    return ok != 0;
  }

  __device__ bool enqueue_body(uint64_t v, Combine* ct) {  // :527-536
    if (!room_to_enqueue()) return false;  // ring mode: full
    for (int p = (int)ld_const(&q->g).patience; p >= 0; --p) {
      uint64_t cid = 0;
      if (enq_fast(v, &cid, ct)) return true;
      if (cid == kWithdrawn) return false;  // pool exhausted
      if (p == 0) return enq_slow(v, cid);
    }
    return false;
  }
  __device__ bool enqueue(uint64_t v, Combine* ct = nullptr) {
    prologue();
    const bool ok = enqueue_body(v, ct);
    epilogue();
    return ok;
  }
  __device__ uint64_t dequeue() {
    prologue();
    const uint64_t v = dequeue_body();
    epilogue();
    return v;
  }

  __device__ uint64_t dequeue_body() {  // :538-555
    // Every ticket comes from claim_h, which takes no cell when T <= H: then
    // every enqueue cell is already owned by a dequeuer (EMPTY at the T
    // read).  Idle consumers would otherwise burn a cell per attempt.
    uint64_t v = kRetry, cid = 0;
    for (int p = (int)ld_const(&q->g).patience; p >= 0; --p) {
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
  Ops<Probe> o{q, hs, &hs[tid], tid, ld_const(&q->g).cells, pr ? pr : (Probe*)&np};
  return o.enqueue(v, ct);
}
template <class Probe = NoProbe>
__device__ __forceinline__ uint64_t dequeue(Queue* q, Handle* hs, uint32_t tid, Probe* pr = nullptr) {
  NoProbe np;
  Ops<Probe> o{q, hs, &hs[tid], tid, ld_const(&q->g).cells, pr ? pr : (Probe*)&np};
  return o.dequeue();
}
// Ring mode: every wave of a kernel that operates on the queue calls
// begin() before its first operation and end() after its last (any active
// lanes; lanes that run no operation may call them too).  begin() marks the
// wave's floor unpublished in LDS; end() withdraws a published floor, so a
// wave that stops using the queue does not hold reclamation back.
__device__ __forceinline__ void begin(Queue* q) {
  if (!ld_const(&q->g).ring) return;
  if (lane_id() == first_lane(active_mask())) wave_lds()->pub[wave_in_block()] = 0;
}
__device__ __forceinline__ void end(Queue* q, Handle* hs, uint32_t tid) {
  if (!ld_const(&q->g).ring) return;
  if (lane_id() == first_lane(active_mask())) {
    WaveLds* s = wave_lds();
    const uint32_t wv = wave_in_block();
    if (s->pub[wv] & 1u) st_rlx(&wave_lo(q, hs)[tid / kWave], kIdle);
    s->pub[wv] = 0;
  }
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
  const uint64_t hw = min(high_water(q), ld_const(&q->g).cells);
  uint64_t* const vw = vals(q);
  uint64_t* const rw = (uint64_t*)refs(q);
  for (uint64_t i = gid; i < hw; i += stride) {
    st_rlx(&vw[i], 0ull);
    st_rlx(&rw[i], 0ull);
  }
}
__device__ inline void reset_handle(Queue* q, Handle* hs, uint32_t tid) {
  const uint32_t n = ld_const(&q->g).nthreads;
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
