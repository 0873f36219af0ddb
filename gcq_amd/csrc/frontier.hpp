// frontier.hpp — frontier queues for the BFS kernels (single, batched and
// served), shared by bfs.hip and batch.hip.
//
// A frontier adapter wraps one queue of any of the four kinds and carries a
// 32-bit payload in the queue itself (G-WFQ / G-LFQ ring indices, SFQ items,
// WFQ values + 1): one queue operation per frontier vertex instead of the
// reference's data-array indirection (bfs_bench.cpp:35-132 goes through
// Enqueue_Ptr / fq_enqueue_ptr: two ring operations and a fence).
// Capacity is sized by the host to the largest frontier, so pushes do not
// fail (the reference ignores failed pushes, bfs_bench.cpp:380, 404, 475).
#pragma once
#include <type_traits>
#include <algorithm>

#include "queue_impl.hpp"

namespace gcqf {

using gcq::lane_id;

constexpr uint64_t kBarrierSpinLimit = 1ull << 25;
// A level's starting Head when it is not known (see kHeadFromDrain below).
constexpr uint64_t kNoHead = ~0ull;

// ---------------------------------------------------------------- adapters
// Counted pops (kCounted): in a level-synchronous traversal a level's input
// ring holds exactly the tickets [Head, Tail) at the level's start, all of
// them written before the previous barrier.  A pop is then one aggregated FAA
// on Head compared with that end: no Head/Tail claim reads, and the level's
// last pop costs one FAA.  Head overshoots the end by up to a wave's lanes per
// wave; drained() at the barrier sets it back to the end.
// kHeadFromDrain: drained() leaves Head equal to the level's end, so a ring's
// Head at the start of the next level it feeds is that end, known to every
// wave (two frontiers alternate, so it is the end of two levels before).
// Such a ring also offers peek(T) / take(T, end, word): when a level's
// tickets [h0, end) fit in the grid's lanes, each lane takes one ticket of
// [h0, h0 + lanes) fixed by its grid position, without an FAA on Head (each
// ticket is taken once, as the FAAs would have handed them out), and its
// slot read goes out together with the read of the level's end.  Head stays
// at h0 and is set to the end by drained(), as after FAA pops.
struct FGwf {
  static constexpr int kQueues = 2;
  static constexpr bool kCounted = true;
  static constexpr bool kHeadFromDrain = true;
  gcq::gwf::Ring* r;
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ uint64_t* o_tail() const { return (uint64_t*)((char*)r + gcq::ld_const(&r->g).off_tail); }
  __device__ __forceinline__ void rewind() const {}
  // Head and Tail words pack (count << thr_bits | thread index); the index
  // field is the null index whenever no slow-path request is being
  // published, as at a level's start and end.
  __device__ __forceinline__ uint64_t level_end() const { return gcq::ld_rlx(o_tail()) >> gcq::ld_const(&r->g).thr_bits; }
  __device__ __forceinline__ int pop_counted(uint32_t tid, uint64_t end, uint32_t* v) const {
    using namespace gcq::gwf;
    NoProbe np;
    RingOps<NoProbe, true> o(r, tid, &np);
    const uint64_t T = o.cnt_of(gcq::wave_faa(o.head_(r), (uint64_t)1 << o.g.thr_bits));
    if (T >= end) return -1;
    return take(tid, T, end, peek(T), v);
  }
  __device__ __forceinline__ uint64_t peek(uint64_t T) const {
    gcq::gwf::NoProbe np;
    gcq::gwf::RingOps<gcq::gwf::NoProbe, true> o(r, 0, &np);
    return gcq::ld_rlx(&gcq::gwf::slots(r)[o.slot(T)]);
  }
  // Ticket T with its slot word E as first read: as pop_counted after its FAA.
  __device__ __forceinline__ int take(uint32_t tid, uint64_t T, uint64_t end, uint64_t E, uint32_t* v) const {
    using namespace gcq::gwf;
    if (T >= end) return -1;
    NoProbe np;
    RingOps<NoProbe, true> o(r, tid, &np);
    const uint32_t j = o.slot(T), tcyc = o.cyc(T);
    for (uint32_t w = 0; (e_cyc(E) != tcyc || e_idx(E) >= kBot) && w < gcq::kSlotWaitSpins; ++w) {
      __builtin_amdgcn_s_sleep(1);
      E = gcq::ld_rlx(&slots(r)[j]);
    }
    if (e_cyc(E) != tcyc || e_idx(E) >= kBot) return 0;
    o.consume(T, j, E);  // also finalises an item a slow-path enqueue left
    *v = e_idx(E);
    return 1;
  }
  // end: the level's end (level_end() at its start): the input ring's Tail,
  // which no one moves during the level, so the barrier need not read it.
  __device__ __forceinline__ void drained(uint64_t end) const {
    const uint32_t B = gcq::ld_const(&r->g).thr_bits;
    gcq::st_rlx((uint64_t*)((char*)r + gcq::ld_const(&r->g).off_head), (end << B) | ((1ull << B) - 1ull));
  }
  __device__ __forceinline__ bool push(uint32_t tid, uint32_t v) const {
    gcq::gwf::NoProbe np;
    gcq::gwf::RingOps<gcq::gwf::NoProbe, true> o(r, tid, &np);
    o.enqueue(v);
    return true;
  }
  __device__ __forceinline__ bool pop(uint32_t tid, uint32_t* v) const {
    gcq::gwf::NoProbe np;
    gcq::gwf::RingOps<gcq::gwf::NoProbe, true> o(r, tid, &np);
    const uint32_t x = o.dequeue(true);
    if (x == gcq::gwf::kBot) return false;
    *v = x;
    return true;
  }
};
struct FLfq {
  static constexpr int kQueues = 2;
  static constexpr bool kCounted = true;
  static constexpr bool kHeadFromDrain = true;
  gcq::lfq::Ring* r;
  // Set by level-synchronous callers whose pops are all counted (they take
  // only tickets pushed before the level's barrier): a push then stores a
  // first-lap ticket's entry instead of CASing it.
  uint32_t counted_push = 0;
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ void rewind() const {}
  // End of the level's tickets: the input ring's Tail, which no one moves
  // during the level.
  __device__ __forceinline__ uint64_t level_end() const { return gcq::ld_rlx(gcq::lfq::tail_p<true>(r)); }
  // 1 = *v popped, 0 = this ticket holds no item (skip it), -1 = the level
  // is drained.
  __device__ __forceinline__ int pop_counted(uint32_t tid, uint64_t end, uint32_t* v) const {
    const uint64_t T = gcq::wave_faa(gcq::lfq::head_p<true>(r), (uint64_t)1);
    if (T >= end) return -1;
    return take(tid, T, end, peek(T), v);
  }
  __device__ __forceinline__ uint64_t peek(uint64_t T) const {
    using namespace gcq::lfq;
    const Geometry g = gcq::ld_const(&r->g);
    return gcq::ld_rlx(slots(r) + slot_of(T, g.slots_log2, g.remap_log2));
  }
  // Ticket T with its slot word E as first read: as pop_counted after its FAA.
  __device__ __forceinline__ int take(uint32_t, uint64_t T, uint64_t end, uint64_t E, uint32_t* v) const {
    using namespace gcq::lfq;
    if (T >= end) return -1;
    const Geometry g = gcq::ld_const(&r->g);
    const uint32_t tcyc = cycle_of(T, g.slots_log2);
    uint64_t* const e = slots(r) + slot_of(T, g.slots_log2, g.remap_log2);
    // Written before the barrier; the wait only guards against a ticket an
    // enqueuer abandoned.
    for (uint32_t w = 0; (e_cyc(E) != tcyc || e_idx(E) >= kBot) && w < gcq::kSlotWaitSpins; ++w) {
      __builtin_amdgcn_s_sleep(1);
      E = gcq::ld_rlx(e);
    }
    if (e_cyc(E) != tcyc || e_idx(E) >= kBot) return 0;
    gcq::or_rlx(e, (uint64_t)kBotC);  // consume: index := ⊥c, keep cycle/safe
    *v = e_idx(E);
    return 1;
  }
  __device__ __forceinline__ void drained(uint64_t end) const { gcq::st_rlx(gcq::lfq::head_p<true>(r), end); }
  __device__ __forceinline__ bool push(uint32_t, uint32_t v) const {
    if (counted_push) gcq::lfq::enqueue<gcq::lfq::NoProbe, true, true>(r, v);
    else gcq::lfq::enqueue<gcq::lfq::NoProbe, true>(r, v);
    return true;
  }
  __device__ __forceinline__ bool pop(uint32_t, uint32_t* v) const {
    const uint32_t x = gcq::lfq::dequeue<gcq::lfq::NoProbe, true>(r, nullptr, true);
    if (x == gcq::lfq::kBot) return false;
    *v = x;
    return true;
  }
};
// SFQ's ticket/phase protocol on slots that carry a 32-bit item: the slot's
// 64-bit word holds the phase in [63:32] and the item in [31:0], so one load
// gives both and one store publishes both.  Ticket t's slot is t mod qlen,
// its lap l = t / qlen; phase 2l (mod 2^32) is empty for the lap and 2l + 1
// full.  Against the queue's general form (item and 64-bit phase in two
// words) a pop saves the item's load and a push the wait between its item
// and phase stores.  The ring is the queue's own (its slots' phase words
// start at 0 = lap 0 empty); the item words stay unused.
struct FSfq {
  static constexpr int kQueues = 2;
  static constexpr bool kCounted = true;
  static constexpr bool kHeadFromDrain = true;
  gcq::sfq::Queue* q;
  // Set by a caller that pushes at most the ring's length into it between
  // resets (one push per vertex into a ring of at least V slots): every
  // ticket is then on lap 0, whose slots init left empty, so a push stores
  // without first reading its slot's phase and a counted pop does not hand
  // its slot on to a next lap.
  uint32_t single_lap = 0;
  __device__ __forceinline__ uint64_t* word(uint64_t t) const {
    return &gcq::sfq::slots(q)[t & (gcq::ld_const(&q->g).qlen - 1)].id;
  }
  __device__ __forceinline__ uint32_t lap_phase(uint64_t t) const { return (uint32_t)(2ull * (t >> gcq::ld_const(&q->g).qlen_log2)); }
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ void rewind() const {}
  // Counted pops: the slot of a ticket below the level's end carries its
  // item and the full phase; the pop takes the item from the word it read
  // and hands the slot to the next lap, as a blocking dequeue does after
  // its wait.
  __device__ __forceinline__ uint64_t level_end() const { return gcq::ld_rlx(gcq::sfq::tail(q)); }
  __device__ __forceinline__ int pop_counted(uint32_t tid, uint64_t end, uint32_t* v) const {
    const uint64_t T = gcq::wave_faa(gcq::sfq::head(q), (uint64_t)1);
    if (T >= end) return -1;
    return take(tid, T, end, peek(T), v);
  }
  __device__ __forceinline__ uint64_t peek(uint64_t T) const { return gcq::ld_rlx(word(T)); }
  // Ticket T with its slot's word as first read: as pop_counted after its FAA.
  __device__ __forceinline__ int take(uint32_t, uint64_t T, uint64_t end, uint64_t first, uint32_t* v) const {
    if (T >= end) return -1;
    const uint32_t full = lap_phase(T) + 1u;
    uint64_t w = first;
    for (uint32_t k = 0; (uint32_t)(w >> 32) != full && k < gcq::kSlotWaitSpins; ++k) {
      __builtin_amdgcn_s_sleep(1);
      w = gcq::ld_rlx(word(T));
    }
    if ((uint32_t)(w >> 32) != full) return 0;
    if (!single_lap) gcq::st_rlx(word(T), (uint64_t)(full + 1u) << 32);  // empty for the next lap (none: single lap)
    *v = (uint32_t)w;
    return 1;
  }
  __device__ __forceinline__ void drained(uint64_t end) const { gcq::st_rlx(gcq::sfq::head(q), end); }
  // Blocking enqueue (sfqueue_hip.cpp:55-70): a ticket by wave-aggregated
  // FAA on Tail, a wait (bounded) for the slot's previous lap to be taken,
  // then item and full phase in one store.
  __device__ __forceinline__ bool push(uint32_t, uint32_t v) const {
    const uint64_t t = gcq::wave_faa(gcq::sfq::tail(q), (uint64_t)1);
    const uint32_t empty = lap_phase(t);
    for (uint64_t k = 0; !single_lap && (uint32_t)(gcq::ld_rlx(word(t)) >> 32) != empty; ++k) {
      if (k > (1ull << 22)) return false;
      gcq::sfq::slot_wait(k, gcq::ld_const(&q->g).backoff);
    }
    gcq::st_rlx(word(t), ((uint64_t)(empty + 1u) << 32) | v);
    return true;
  }
  // Non-blocking pop (host-loop and barrier-free modes, the ray tracer): the
  // queue's non-blocking dequeue (sfqueue_hip.cpp:109-128) on packed slots.
  // The wave reads Head once, tests its lanes' tickets, and claims the
  // prefix whose slots are full with one CAS; it answers EMPTY only after
  // reading Head >= Tail (head read first), and retries BUSY otherwise.
  __device__ bool pop(uint32_t, uint32_t* v) const {
    uint64_t* const hp = gcq::sfq::head(q);
    for (uint32_t r = 0; r < (1u << 20); ++r) {
      const uint64_t m = gcq::active_mask();
      const uint32_t leader = gcq::first_lane(m);
      uint64_t t0 = 0;
      if (gcq::lane_id() == leader) t0 = gcq::ld_rlx(hp);
      t0 = gcq::bcast64(t0, leader);
      const uint64_t t = t0 + gcq::rank_in(m);
      const uint64_t w = gcq::ld_rlx(word(t));
      const bool ready = (uint32_t)(w >> 32) == lap_phase(t) + 1u;
      const uint64_t not_ready = m & ~__ballot(ready);
      const uint32_t k = not_ready ? (uint32_t)__builtin_popcountll(m & ((1ull << gcq::first_lane(not_ready)) - 1ull))
                                   : (uint32_t)__builtin_popcountll(m);
      uint32_t won = 0;
      if (gcq::lane_id() == leader && k > 0) {
        uint64_t exp = t0;
        won = gcq::cas_rlx(hp, exp, t0 + k) ? 1u : 0u;
      }
      won = gcq::bcast32(won, leader);
      if (won && gcq::rank_in(m) < k) {
        gcq::st_rlx(word(t), (uint64_t)(lap_phase(t) + 2u) << 32);  // empty for the next lap
        *v = (uint32_t)w;
        return true;
      }
      if (won) continue;  // lanes past the claimed prefix try again at once
      uint32_t empty = 0;
      if (gcq::lane_id() == leader) {
        const uint64_t h = gcq::ld_rlx(hp);
        gcq::vm_wait();  // Head is read before Tail: Head (at the Tail read) >= h >= Tail means empty
        empty = h >= gcq::ld_rlx(gcq::sfq::tail(q)) ? 1u : 0u;
      }
      if (gcq::bcast32(empty, leader)) return false;
      gcq::sfq::backoff(8);
    }
    return false;
  }
};
// WFQ cells are never reused by the algorithm, and every level burns up to one
// cell per idle lane, so the kernels rotate three WFQ frontiers: while one is
// the level's input and one its output, the third (drained a level earlier)
// is cleared by the grid and rewound at the level's barrier.
struct FYmc {  // values are v + 1: 0 is the queue's empty cell (wfqueue_hip_opt.hpp:42-46)
  static constexpr int kQueues = 3;
  static constexpr bool kCounted = true;
  static constexpr bool kHeadFromDrain = false;  // rewound a level earlier instead
  gcq::ymc::Queue* q;
  gcq::ymc::Handle* h;
  // Counted pops: a level's input queue was only enqueued into during the
  // level before, so every cell in [H, T) holds its value (an enqueue's CAS
  // fails only on a cell a dequeuer marked).  A pop reads the value; no
  // dequeue request is pending, so no helper competes for the cell.  The
  // queue is recycled and rewound as the idle frontier a level later, which
  // also clears Head's overshoot.
  __device__ __forceinline__ uint64_t level_end() const { return gcq::ld_rlx(gcq::ymc::Tp(q)); }
  __device__ __forceinline__ int pop_counted(uint32_t, uint64_t end, uint32_t* v) const {
    using namespace gcq::ymc;
    const uint64_t T = gcq::wave_faa(Hp(q), (uint64_t)1);
    if (T >= end) return -1;
    if (T >= gcq::ld_const(&q->g).cells) return 0;
    uint64_t* const c = &vals(q)[T];
    uint64_t x = gcq::ld_rlx(c);
    for (uint32_t w = 0; x == kBottom && w < gcq::kSlotWaitSpins; ++w) {
      __builtin_amdgcn_s_sleep(1);
      x = gcq::ld_rlx(c);
    }
    if (x == kBottom || x == kTop || x == kRetry) return 0;
    *v = (uint32_t)(x - 1);
    return 1;
  }
  __device__ __forceinline__ void drained(uint64_t) const {}
  __device__ __forceinline__ void recycle(uint64_t gid, uint64_t stride, uint32_t tid) const {
    gcq::ymc::recycle_cells(q, gid, stride);
    gcq::ymc::reset_handle(q, h, tid);
  }
  __device__ __forceinline__ void rewind() const { gcq::ymc::rewind(q); }
  __device__ __forceinline__ bool push(uint32_t tid, uint32_t v) const {
    return gcq::ymc::enqueue(q, h, tid, (uint64_t)v + 1);
  }
  __device__ __forceinline__ bool pop(uint32_t tid, uint32_t* v) const {
    const uint64_t x = gcq::ymc::dequeue(q, h, tid);
    if (x == gcq::ymc::kEmpty) return false;
    *v = (uint32_t)(x - 1);
    return true;
  }
};

// Frontier queues of one traversal: level L reads q[L % K] and writes
// q[(L + 1) % K]; with K = 3 the third one is recycled during the level.
template <class F>
struct Frontiers {
  F q[3];
  __device__ __forceinline__ const F& in(int32_t L) const { return q[L % F::kQueues]; }
  __device__ __forceinline__ const F& out(int32_t L) const { return q[(L + 1) % F::kQueues]; }
  __device__ __forceinline__ const F& idle(int32_t L) const { return q[(L + 2) % 3]; }
};

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Enqueue the first n staged items (n <= 64) of a wave's LDS stage and
// shift the rest down; `sc` (wave-uniform) is the staged count.
template <class F>
__device__ __forceinline__ void flush(const F& qout, uint32_t tid, uint32_t* stage, uint32_t& sc, uint32_t n,
                                      uint32_t* err) {
  const uint32_t lane = lane_id();
  const uint32_t v = stage[lane];
  if (lane < n && !qout.push(tid, v)) gcq::or_rlx(err, 1u);
  const uint32_t rest = sc - n;
  const uint32_t mv = lane < rest ? stage[64 + lane] : 0u;
  if (lane < rest) stage[lane] = mv;
  sc = rest;
}

// Barrier words, each on its own 128-B line.
struct Barrier {
  uint32_t count, p0[31];
  uint32_t gen, flag, p1[30];  // flag: what the last block's last() returned, stored before gen
  uint32_t abort, p2[31];
};

// Sense-reversing grid barrier over co-resident blocks.  The last block to
// arrive runs `last()` before releasing the others; when last() returns a
// value, every thread of every block gets it in *flag (one read per block,
// after the release).  Returns false once any block has aborted (spin
// limit), so every block leaves its loop.
template <class Last>
__device__ bool grid_barrier(Barrier* b, uint32_t nblocks, Last last, uint32_t* flag = nullptr) {
  __shared__ uint32_t s_flag, s_ok;
  gcq::vm_wait();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = gcq::ld_rlx(&b->gen);
    gcq::vm_wait();  // generation read before arriving
    const uint32_t a = gcq::faa_rlx(&b->count, 1u);
    uint32_t f = 0;
    if (a == nblocks - 1) {
      if constexpr (std::is_void_v<decltype(last())>) {
        last();
      } else {
        f = last();
        gcq::st_rlx(&b->flag, f);
      }
      gcq::st_rlx(&b->count, 0u);
      gcq::vm_wait();
      gcq::st_rlx(&b->gen, g + 1u);
    } else {
      uint64_t spins = 0;
      while (gcq::ld_rlx(&b->gen) == g) {
        if (gcq::ld_rlx(&b->abort)) break;
        if (++spins > kBarrierSpinLimit) {
          gcq::st_rlx(&b->abort, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if constexpr (!std::is_void_v<decltype(last())>) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the flag stored before the release
        f = gcq::ld_rlx(&b->flag);
      }
    }
    s_flag = f;
    s_ok = gcq::ld_rlx(&b->abort) == 0u;  // one read per block, not one per wave
  }
  __syncthreads();
  if (flag) *flag = s_flag;
  return s_ok != 0u;
}

// ---------------------------------------------------------------- hub vertices
// A vertex whose adjacency is long (R-MAT hubs reach 162,818 edges at
// scale 22) would hold its level open while one wave scans it 64 edges a
// round.  A wave that settles such a vertex posts it instead, and after the
// level barrier the whole grid scans the posted edges in chunks, then meets
// at a second barrier; a level that posts nothing keeps its one barrier.
// (Enterprise's split of the frontier by degree, Liu & Huang, SC'15, with
// the grid taking the largest class.)  Pools alternate by level parity: a
// level's pool is read by every block after the level's first barrier and
// cleared at the end of the next level, when no block can still read it.
constexpr int32_t kHubDeg = 4096;     // more edges than this: posted
constexpr uint32_t kHubChunk = 1024;  // edges per claimed chunk
constexpr uint32_t kMaxHubs = 2048;   // entries per level; a vertex past them is scanned by its wave
constexpr uint32_t kHubCursors = 8;   // chunk claims striped over 8 lines

struct HubEntry {
  uint32_t item, base;  // (group, vertex); its first chunk
  int32_t beg, deg;     // its edge range
  uint64_t fresh;       // the members sent along its edges
};
struct HubPool {
  uint64_t word, p0[15];  // [63:40] entries, [39:0] chunks, posted this level
  uint32_t cursor[kHubCursors][32];
  HubEntry e[kMaxHubs];
};
constexpr uint64_t kHubHeaderWords = (128 + 128 * kHubCursors) / 8;  // what a reset clears
static_assert(sizeof(HubPool) == 8 * kHubHeaderWords + sizeof(HubEntry) * kMaxHubs, "hub pool");

// Post one vertex; false when the level's table is full.  `fresh` is the
// multi-source form's member mask (unused by the single-source form).
__device__ __forceinline__ bool hub_post(HubPool* hp, uint32_t item, int32_t beg, int32_t deg, uint64_t fresh) {
  const uint64_t nch = ((uint64_t)deg + kHubChunk - 1) / kHubChunk;
  uint64_t w = gcq::ld_rlx(&hp->word);
  for (;;) {
    if ((w >> 40) >= kMaxHubs) return false;
    if (gcq::cas_rlx(&hp->word, w, w + (1ull << 40) + nch)) break;
  }
  HubEntry* e = &hp->e[w >> 40];
  gcq::st_rlx(&e->item, item);
  gcq::st_rlx(&e->base, (uint32_t)(w & ((1ull << 40) - 1)));
  gcq::st_rlx(&e->beg, beg);
  gcq::st_rlx(&e->deg, deg);
  gcq::st_rlx(&e->fresh, fresh);
  return true;
}

// Level end: clear the other parity's pool (last used a level ago, next
// used a level from now).  Run by one thread, before a barrier's release.
__device__ __forceinline__ void hub_reset(HubPool* hp) {
  gcq::st_rlx(&hp->word, 0ull);
  for (uint32_t c = 0; c < kHubCursors; ++c) gcq::st_rlx(&hp->cursor[c][0], 0u);
}

// The grid's scan of a level's posted vertices, between the level's two
// barriers (all threads of every block call it; the pool is final).  Wave
// w < P (P = min(chunks, waves)) first takes chunk w, then chunks
// P + s + 8k by fetch-add on cursor s, moving to the next cursor when one
// runs out.  scan(entry, first, stop) runs wave-wide over the entry's edges
// [first, stop) (offsets into its adjacency).
template <class Scan>
__device__ void hub_scan(HubPool* hp, uint32_t tid, uint32_t nwaves, Scan&& scan) {
  __shared__ uint32_t s_base[kMaxHubs];
  const uint64_t word = gcq::ld_rlx(&hp->word);
  const uint32_t n = (uint32_t)(word >> 40);
  const uint64_t total = word & ((1ull << 40) - 1);
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_base[i] = gcq::ld_rlx(&hp->e[i].base);
  __syncthreads();
  const uint32_t wv = tid / 64;
  const uint64_t P = total < nwaves ? total : nwaves;
  if (wv >= P) return;
  uint64_t c = wv;
  uint32_t s = wv % kHubCursors, tried = 0;
  for (;;) {
    while (c >= total && tried < kHubCursors) {
      uint32_t k = 0;
      if (lane_id() == 0) k = gcq::faa_rlx(&hp->cursor[s][0], 1u);
      c = P + s + (uint64_t)kHubCursors * __shfl(k, 0, 64);
      if (c >= total) {
        s = (s + 1) % kHubCursors;
        ++tried;
      }
    }
    if (c >= total) break;
    uint32_t lo = 0, hi = n;  // last entry whose first chunk is at or below c
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (s_base[mid] <= c) lo = mid;
      else hi = mid;
    }
    const HubEntry* e = &hp->e[lo];
    const int32_t deg = gcq::ld_cached(&e->deg);
    const int32_t first = (int32_t)(c - s_base[lo]) * (int32_t)kHubChunk;
    scan(e, first, first + (int32_t)kHubChunk < deg ? first + (int32_t)kHubChunk : deg);
    c = total;  // the next chunk by claim
  }
}

// Level barrier of the persistent BFS: arrival and the level's discovery
// count travel in one 64-bit FAA per block ([63:32] blocks arrived, [31:0]
// vertices discovered), and the release word carries the next level's size
// ([63:32] level + 1, [31:0] count).  One atomic and a spin per level and
// block, instead of a counter FAA per wave, a generation read, the arrival
// FAA and a count read.
struct LevelBarrier {
  uint64_t arrive, p0[15];
  uint64_t release, p1[15];
  uint32_t abort, p2[31];
};

// All threads of the block call it, with seq = the number of level
// barriers the kernel has passed; wave_disc is the calling wave's discovery
// count (wave-uniform).  *next gets the discoveries since the last barrier
// and *cum (per thread, the same in every thread) the running total.  The
// arrival word only grows: [63:32] counts arrivals over all barriers and
// [31:0] the discoveries over all levels (below 2^31: each vertex is
// discovered once), so the last block needs no reset store.  The last block
// to arrive runs last() first: kBarrierStores in its answer means it made
// stores that must land before the release, kBarrierHubs that the level
// posted long adjacencies (bit 31 of the release word carries it to *hubs).
// Returns false once any block has aborted (spin limit).
constexpr uint32_t kBarrierStores = 1, kBarrierHubs = 2;
template <class Last>
__device__ bool level_barrier(LevelBarrier* b, uint32_t nblocks, uint32_t seq, uint32_t wave_disc, uint32_t* next,
                              uint32_t* cum, Last last, bool* hubs = nullptr) {
  __shared__ uint32_t s_wd[16];
  __shared__ uint32_t s_next, s_ok;
  gcq::vm_wait();  // this wave's pushes have completed
  if (lane_id() == 0) s_wd[(threadIdx.x / 64) & 15] = wave_disc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (uint32_t k = 0; k < (blockDim.x + 63) / 64; ++k) sum += s_wd[k];
    const uint64_t old = gcq::faa_rlx(&b->arrive, (1ull << 32) | sum);
    uint32_t ok = 1, n = 0;
    if ((uint32_t)(old >> 32) == (seq + 1u) * nblocks - 1u) {
      n = (uint32_t)old + sum - *cum;
      const uint32_t f = last();
      if (f & kBarrierStores) gcq::vm_wait();  // last()'s stores complete before the release
      if (f & kBarrierHubs) n |= 1u << 31;
      gcq::st_rlx(&b->release, ((uint64_t)(seq + 1u) << 32) | n);
    } else {
      for (uint64_t spins = 0;; ++spins) {
        const uint64_t r = gcq::ld_rlx(&b->release);
        if ((uint32_t)(r >> 32) == seq + 1u) {
          n = (uint32_t)r;
          break;
        }
        if (gcq::ld_rlx(&b->abort) || spins > kBarrierSpinLimit) {
          gcq::st_rlx(&b->abort, 1u);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    s_next = n;
    s_ok = ok;
  }
  __syncthreads();
  *next = s_next & 0x7FFFFFFFu;
  *cum += *next;
  if (hubs) *hubs = (s_next >> 31) != 0u;
  return s_ok != 0;
}

// ---------------------------------------------------------------- reset kernels
namespace {
__global__ void kf_fill_i32(int32_t* p, uint64_t n, int32_t v) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < n; i += st) p[i] = v;
}
__global__ void kf_zero_u64(uint64_t* p, uint64_t n) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < n; i += st) p[i] = 0;
}
__global__ void kf_reset_gwf(void* a, void* b, gcq::gwf::InitParams p) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::gwf::init_slice((gcq::gwf::Ring*)a, p, false, gid, st);
  gcq::gwf::init_slice((gcq::gwf::Ring*)b, p, false, gid, st);
}
__global__ void kf_reset_lfq(void* a, void* b, uint32_t n) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::lfq::init_slice((gcq::lfq::Ring*)a, n, false, 0, gid, st, 0);
  gcq::lfq::init_slice((gcq::lfq::Ring*)b, n, false, 0, gid, st, 0);
}
__global__ void kf_reset_sfq(void* a, void* b, uint64_t n) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::sfq::init_slice((gcq::sfq::Queue*)a, n, 0, gid, st);
  gcq::sfq::init_slice((gcq::sfq::Queue*)b, n, 0, gid, st);
}
__global__ void kf_reset_ymc(void* q, void* h, uint64_t cells, uint32_t nthreads) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::ymc::init_slice((gcq::ymc::Queue*)q, (gcq::ymc::Handle*)h, cells, nthreads, 16, gid, st);
}
}  // namespace

inline uint32_t grid_of(uint64_t work) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, 8192));
}
inline uint32_t pow2_at_least(uint64_t x) {
  uint64_t p = 64;
  while (p < x) p <<= 1;
  return (uint32_t)std::min<uint64_t>(p, 1ull << 31);
}

// ---------------------------------------------------------------- host side
// The frontier queues of one traversal context, in HBM.
struct FrontierSet {
  int kind = -1, nq = 0;
  uint32_t threads = 0, capacity = 0;
  uint64_t ymc_cells = 0;
  void* q[3] = {nullptr, nullptr, nullptr};
  void* h[3] = {nullptr, nullptr, nullptr};  // WFQ handles

  void release() {
    for (int i = 0; i < 3; ++i) {
      if (q[i]) (void)hipFree(q[i]);
      if (h[i]) (void)hipFree(h[i]);
    }
    *this = FrontierSet();
  }

  // Queues for up to `items` frontier entries per level, used by at most
  // `nthreads` threads.  Reuses the allocation when it is large enough.
  int ensure(int k, uint64_t items, uint32_t nthreads) {
    const uint32_t cap = pow2_at_least(std::max<uint64_t>(items, nthreads));
    if (k == kind && nthreads <= threads && cap <= capacity && q[0]) return GCQ_OK;
    release();
    uint64_t qbytes = 0;
    switch (k) {
      case GCQ_GWFQ:
        GCQ_REQUIRE(cap <= (1u << 25), "G-WFQ frontier capacity %u exceeds 2^25 entries", cap);
        qbytes = gcq::gwf::ring_bytes(cap, nthreads);
        break;
      case GCQ_GLFQ: qbytes = gcq::lfq::ring_bytes(cap); break;
      case GCQ_SFQ: qbytes = gcq::sfq::queue_bytes(cap); break;
      case GCQ_WFQ:
        // One level's cells: its frontier plus at most one burned cell per
        // dequeuing lane; the queue is recycled two levels later.
        ymc_cells = (uint64_t)cap + 4ull * nthreads + (1ull << 20);
        qbytes = gcq::ymc::queue_bytes(ymc_cells);
        break;
      default: GCQ_REQUIRE(false, "unknown queue kind %d", k);
    }
    nq = k == GCQ_WFQ ? 3 : 2;
    for (int i = 0; i < nq; ++i) {
      if (hipMalloc(&q[i], qbytes) != hipSuccess ||
          (k == GCQ_WFQ && hipMalloc(&h[i], sizeof(gcq::ymc::Handle) * (size_t)nthreads) != hipSuccess)) {
        release();
        gcqh::set_error("out of device memory for frontier queues (%llu bytes each)", (unsigned long long)qbytes);
        return GCQ_ENOMEM;
      }
    }
    kind = k;
    threads = nthreads;
    capacity = cap;
    return GCQ_OK;
  }

  // Empty every queue (stream-ordered kernels; no memset nodes, so the
  // sequence may be captured into a graph).
  void reset(hipStream_t s) const {
    switch (kind) {
      case GCQ_GWFQ: {
        gcq::gwf::InitParams p{capacity, threads, 0, 64, 128, 256, 0};
        hipLaunchKernelGGL(kf_reset_gwf, dim3(grid_of(2ull * capacity)), dim3(256), 0, s, q[0], q[1], p);
        break;
      }
      case GCQ_GLFQ:
        hipLaunchKernelGGL(kf_reset_lfq, dim3(grid_of(2ull * capacity)), dim3(256), 0, s, q[0], q[1], capacity);
        break;
      case GCQ_SFQ:
        hipLaunchKernelGGL(kf_reset_sfq, dim3(grid_of(capacity)), dim3(256), 0, s, q[0], q[1], (uint64_t)capacity);
        break;
      case GCQ_WFQ:
        for (int i = 0; i < 3; ++i) {
          hipLaunchKernelGGL(kf_zero_u64, dim3(grid_of(ymc_cells * 2)), dim3(256), 0, s,
                             (uint64_t*)((char*)q[i] + gcq::ymc::kHdrBytes), ymc_cells * 2);
          hipLaunchKernelGGL(kf_reset_ymc, dim3(grid_of(threads)), dim3(256), 0, s, q[i], h[i], ymc_cells, threads);
        }
        break;
    }
  }

  // This set's queues through adapter F (which must match its kind).
  template <class F>
  Frontiers<F> view() const {
    Frontiers<F> f;
    for (int i = 0; i < 3; ++i) {
      if constexpr (std::is_same_v<F, FGwf>) f.q[i] = FGwf{(gcq::gwf::Ring*)q[i % 2]};
      else if constexpr (std::is_same_v<F, FLfq>) f.q[i] = FLfq{(gcq::lfq::Ring*)q[i % 2], 0u};
      else if constexpr (std::is_same_v<F, FSfq>) f.q[i] = FSfq{(gcq::sfq::Queue*)q[i % 2], 0u};
      else f.q[i] = FYmc{(gcq::ymc::Queue*)q[i], (gcq::ymc::Handle*)h[i]};
    }
    return f;
  }

  // Call fn(Frontiers<F>) with the adapter type of this set's kind.
  template <class Fn>
  int dispatch(Fn&& fn) const {
    switch (kind) {
      case GCQ_GWFQ: return fn(view<FGwf>());
      case GCQ_GLFQ: return fn(view<FLfq>());
      case GCQ_SFQ: return fn(view<FSfq>());
      case GCQ_WFQ: return fn(view<FYmc>());
    }
    GCQ_REQUIRE(false, "frontier set not allocated");
  }
};

}  // namespace gcqf
