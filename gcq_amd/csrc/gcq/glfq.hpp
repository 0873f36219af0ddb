// glfq.hpp — G-LFQ: lock-free bounded MPMC ring (SCQ) with one aggregated
// ticket FAA per wave, plus the two-ring index indirection for u64 payloads.
//
// Algorithm follows the reference glf_queue.hpp (SCQ, Nikolaev DISC'19):
//   entry  = Cycle(8)@48 | IsSafe@47 | Index(32)       glf_queue.hpp:56-90
//   enqueue / dequeue / catchup                        glf_queue.hpp:447-560
//   indirection enqueue_ptr / dequeue_ptr              glf_queue.hpp:566-582
// MI355X layout (own design): the ring header is one 2 KiB block whose
// 128-byte lines hold the immutable geometry and, on separate lines, Tail,
// Head and Threshold (the reference packs all three with Entry[0..4] into one
// line).  Slots follow the header.  Capacity is a runtime field so one binary
// serves every capacity.
#pragma once
#include "prim.hpp"

namespace gcq {
namespace lfq {

constexpr uint32_t kBot = 0xFFFFFFFEu;   // empty slot
constexpr uint32_t kBotC = 0xFFFFFFFFu;  // consumed slot
constexpr int kCycShift = 48;
constexpr int kSafeShift = 47;
constexpr uint64_t kSafe = 1ull << kSafeShift;
constexpr uint64_t kIdxMask = 0xFFFFFFFFull;
constexpr uint32_t kHdrBytes = 2048;

struct Geometry {
  uint32_t n;            // capacity, power of two
  uint32_t slots_log2;   // log2(2n)
  uint32_t remap_log2;   // log2(slots per 128-B line) when remapping, else 0
  int32_t thresh_max;    // 3n - 1
  uint32_t off_tail, off_head, off_thr;  // byte offsets of the counters
  uint32_t magic;
  uint32_t poll_backoff;  // consumer-side polite polling base (0 = off)
};
constexpr uint32_t kMagic = 0x4C465131u;  // "LFQ1"

// Ring = header + 2n slots.  Device code addresses it through this view.
struct Ring {
  Geometry g;
  uint8_t pad_[kHdrBytes - sizeof(Geometry)];
};
static_assert(sizeof(Ring) == kHdrBytes, "ring header size");

__host__ __device__ constexpr uint64_t ring_bytes(uint32_t n) {
  return (uint64_t)kHdrBytes + 2ull * n * 8ull;
}

__device__ __forceinline__ uint64_t* slots(Ring* r) { return (uint64_t*)((char*)r + kHdrBytes); }
__device__ __forceinline__ uint64_t* tail(Ring* r) { return (uint64_t*)((char*)r + r->g.off_tail); }
__device__ __forceinline__ uint64_t* head(Ring* r) { return (uint64_t*)((char*)r + r->g.off_head); }
__device__ __forceinline__ int32_t* thr(Ring* r) { return (int32_t*)((char*)r + r->g.off_thr); }
// The same with the geometry read through the constant address space
// (ld_const) when kC: the BFS frontiers and the ray tracer's tile queues use
// it; the throughput, history and FIFO harnesses keep plain reads (measured:
// constant reads cost the split 25/75 run 15% at 65,536 threads with
// combined claims, while they shorten the frontier and tile queues' chains).
template <bool kC>
__device__ __forceinline__ Geometry geom(Ring* r) {
  if constexpr (kC) return ld_const(&r->g);
  else return r->g;
}
template <bool kC>
__device__ __forceinline__ uint64_t* tail_p(Ring* r) {
  if constexpr (kC) return (uint64_t*)((char*)r + ld_const(&r->g).off_tail);
  else return tail(r);
}
template <bool kC>
__device__ __forceinline__ uint64_t* head_p(Ring* r) {
  if constexpr (kC) return (uint64_t*)((char*)r + ld_const(&r->g).off_head);
  else return head(r);
}
template <bool kC>
__device__ __forceinline__ int32_t* thr_p(Ring* r) {
  if constexpr (kC) return (int32_t*)((char*)r + ld_const(&r->g).off_thr);
  else return thr(r);
}

__host__ __device__ __forceinline__ uint32_t e_cyc(uint64_t e) { return (uint32_t)(e >> kCycShift) & 0xFFu; }
__host__ __device__ __forceinline__ bool e_safe(uint64_t e) { return (e >> kSafeShift) & 1u; }
__host__ __device__ __forceinline__ uint32_t e_idx(uint64_t e) { return (uint32_t)(e & kIdxMask); }
__host__ __device__ __forceinline__ uint64_t e_pack(uint32_t cyc, bool safe, uint32_t idx) {
  return ((uint64_t)(cyc & 0xFFu) << kCycShift) | ((uint64_t)safe << kSafeShift) | (uint64_t)idx;
}
constexpr uint64_t kInit = kSafe | (uint64_t)kBot;  // cycle 0, safe, empty

// Three-way modular cycle order on 8 bits (reference glf_queue.hpp:106-110).
__host__ __device__ __forceinline__ bool cyc_behind(uint32_t a, uint32_t b) {
  return ((a - b) & 0xFFu) >= 0x80u;
}
__host__ __device__ __forceinline__ uint32_t cycle_of(uint64_t t, uint32_t slots_log2) {
  return (uint32_t)(t >> slots_log2) & 0xFFu;
}
// Slot of a ticket.  With remapping, consecutive tickets of one wave land on
// distinct 128-B lines (reference cache_remap, glf_queue.hpp:116-120, used 64-B
// lines; MI355X lines hold 16 slots).
__host__ __device__ __forceinline__ uint32_t slot_of(uint64_t t, uint32_t slots_log2, uint32_t remap_log2) {
  const uint32_t pos = (uint32_t)t & ((1u << slots_log2) - 1u);
  if (remap_log2 == 0) return pos;
  const uint32_t e = 1u << remap_log2;
  return ((pos & (e - 1u)) << (slots_log2 - remap_log2)) | (pos >> remap_log2);
}

// ---------------------------------------------------------------- init
// Grid-stride initialisation: every thread of the launch calls it with its
// own (gid, stride).  A single thread may call it with (0, 1).
__device__ inline void init_slice(Ring* r, uint32_t n, bool full, uint32_t remap_log2, uint32_t gid,
                                  uint32_t stride, uint32_t poll_backoff = 0) {
  const uint32_t slots_log2 = 31u - __builtin_clz(2u * n);
  const uint32_t nslots = 2u * n;
  uint64_t* e = slots(r);
  for (uint32_t pos = gid; pos < nslots; pos += stride) {
    const uint64_t t = (uint64_t)nslots + pos;
    const uint32_t j = slot_of(t, slots_log2, remap_log2);
    e[j] = (full && pos < n) ? e_pack(cycle_of(t, slots_log2), true, pos) : kInit;
  }
  if (gid == 0) {
    Geometry g;
    g.n = n;
    g.slots_log2 = slots_log2;
    g.remap_log2 = remap_log2;
    g.thresh_max = 3 * (int32_t)n - 1;
    g.off_tail = 1 * kLine;
    g.off_head = 5 * kLine;
    g.off_thr = 9 * kLine;
    g.magic = kMagic;
    g.poll_backoff = poll_backoff;
    r->g = g;
    *(uint64_t*)((char*)r + g.off_tail) = (uint64_t)nslots + (full ? n : 0u);
    *(uint64_t*)((char*)r + g.off_head) = (uint64_t)nslots;
    *(int32_t*)((char*)r + g.off_thr) = full ? g.thresh_max : -1;
  }
}

// ---------------------------------------------------------------- ring ops
// Raise Tail to at least h so enqueuers skip tickets dequeuers already
// passed.  The reference retries a CAS from a possibly stale Tail up to 32
// times (glf_queue.hpp:447-455); every failed CAS is another request to the
// hottest line of the ring.  Tail is a plain counter here, so one fetch_max
// has the same effect.
template <bool kC = false>
__device__ __forceinline__ void catchup(Ring* r, uint64_t h) { max_rlx(tail_p<kC>(r), h); }

// Head ticket bounded by the items present (consumer side).  The wave
// leader reads Head, then Tail, and takes k = min(active lanes, Tail - Head)
// tickets with one FAA: consumers never take a ticket no enqueuer has taken
// yet (such a ticket marks its slot and sends the enqueuer that later draws it
// round again).  The reference takes the ticket first (glf_queue.hpp:510-560),
// so tens of thousands of idle consumers drive Head far past Tail.
// Returns this lane's ticket, kNoTicket (none this round) or kEmptyTicket
// (Tail <= Head at the leader's Tail read: EMPTY for the whole wave,
// linearised at that read).
constexpr uint64_t kNoTicket = ~0ull, kEmptyTicket = ~1ull;
template <bool kC = false>
__device__ __forceinline__ uint64_t claim_head(Ring* r) {
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  const uint32_t n = (uint32_t)__builtin_popcountll(m);
  uint64_t base = 0;
  uint32_t k = 0;
  if (lane_id() == leader) {
    const uint64_t h = ld_rlx(head_p<kC>(r));
    vm_wait();  // Head is read before Tail
    const uint64_t t = ld_rlx(tail_p<kC>(r));
    const uint64_t avail = t > h ? t - h : 0;
    k = avail < n ? (uint32_t)avail : n;
    if (k) base = faa_rlx(head_p<kC>(r), (uint64_t)k);
  }
  k = bcast32(k, leader);
  if (k == 0) return kEmptyTicket;
  base = bcast64(base, leader);
  const uint32_t rank = rank_in(m);
  return rank < k ? base + rank : kNoTicket;
}

// claim_head() with the Head/Tail reads and the FAA shared by the waves of
// the workgroup that claim together (combine_claim).
template <bool kC = false>
__device__ __forceinline__ uint64_t claim_head_combined(Ring* r, Combine* cb) {
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  const uint32_t n = (uint32_t)__builtin_popcountll(m);
  uint64_t base = 0;
  uint32_t k = 0, off = 0;
  if (lane_id() == leader) {
    off = combine_claim(cb, n, &k, &base, [&](uint32_t lanes, uint32_t* kk, uint64_t* bb) {
      const uint64_t h = ld_rlx(head_p<kC>(r));
      vm_wait();  // Head is read before Tail
      const uint64_t t = ld_rlx(tail_p<kC>(r));
      const uint64_t avail = t > h ? t - h : 0;
      *kk = avail < lanes ? (uint32_t)avail : lanes;
      *bb = *kk ? faa_rlx(head_p<kC>(r), (uint64_t)*kk) : 0;
    });
  }
  k = bcast32(k, leader);
  if (k == 0) return kEmptyTicket;
  off = bcast32(off, leader);
  base = bcast64(base, leader);
  const uint32_t rank = rank_in(m);
  return off + rank < k ? base + off + rank : kNoTicket;
}

struct NoProbe {
  __device__ __forceinline__ void step() {}
  __device__ __forceinline__ void after_ticket() {}
  __device__ __forceinline__ void slow_entry() {}
};

// Enqueue a slot index; the ring must have room (indirection guarantees it).
// scq_reset: reset the threshold in SCQ's load-then-store form (the
// free-index ring; reset_threshold).  ct: draw the Tail ticket together with
// the workgroup's other enqueuing waves (group_faa).
// kFirstLapStore: the ring's consumers only take tickets whose enqueues
// completed before (level-synchronous frontiers drained by counted pops),
// so no dequeuer marks a slot ahead of its enqueuer; a first-lap ticket's
// slot then still holds its init entry, and a store replaces the CAS.
template <class Probe = NoProbe, bool kC = false, bool kFirstLapStore = false>
__device__ inline void enqueue(Ring* r, uint32_t idx, Probe* pr = nullptr, bool scq_reset = false,
                               Combine* ct = nullptr) {
  const Geometry g = geom<kC>(r);
  uint64_t* const e = slots(r);
  for (;;) {
    const uint64_t T = ct ? group_faa(ct, tail_p<kC>(r), (uint64_t)1) : wave_faa(tail_p<kC>(r), (uint64_t)1);
    // Every earlier memory operation of this wave completes before the slot
    // is touched; this is what orders a payload store before the index is
    // published (and a payload load before the index is recycled).
    vm_wait();
    if (pr) { pr->step(); pr->after_ticket(); }
    const uint32_t tcyc = cycle_of(T, g.slots_log2);
    const uint32_t j = slot_of(T, g.slots_log2, g.remap_log2);
    // Speculative CAS: in steady state the slot holds the previous lap's
    // consumed entry (cycle tcyc - 1, safe, index ⊥c), or the initial entry
    // on the first lap.  When it does, one CAS replaces load + CAS;
    // otherwise the CAS returns the current word and the usual checks run.
    uint64_t E = (T >> g.slots_log2) == 1 ? kInit : e_pack(tcyc - 1u, true, kBotC);
    bool ok;
    if (kFirstLapStore && (T >> g.slots_log2) == 1) {
      st_rlx(&e[j], e_pack(tcyc, true, idx));
      ok = true;
    } else {
      ok = cas_rlx(&e[j], E, e_pack(tcyc, true, idx));  // E refreshed on failure
    }
    while (!ok) {
      if (pr) pr->step();
      if (!(cyc_behind(e_cyc(E), tcyc) && e_idx(E) >= kBot && (e_safe(E) || ld_rlx(head_p<kC>(r)) <= T))) break;
      ok = cas_rlx(&e[j], E, e_pack(tcyc, true, idx));
    }
    // Threshold back to 3n - 1 after the insertions (glf_queue.hpp:493-495):
    // one reset per wave, issued after every successful lane's CAS has
    // returned, instead of a load (and exchange) per lane.  Storing 3n - 1
    // when it already holds 3n - 1 is harmless, and a reset that lands after
    // a concurrent decrement is the same as a late reset.
    const uint64_t okm = __ballot(ok);
    if (okm && lane_id() == first_lane(okm)) reset_threshold(thr_p<kC>(r), g.thresh_max, scq_reset);
    if (ok) return;
  }
}

// Dequeue a slot index, or kBot when the ring is empty.
// idle_check: consumer side, tickets bounded by the items present.  The
// free-index ring of the indirection (idle_check false) takes plain
// aggregated tickets while it has indices and bounded ones for kLowStreak
// dequeues after this wave finds it empty, so producers of a full queue do
// not mark free-ring
// slots ahead of the consumers returning indices.  The per-wave hint only
// picks the ticket path; its initial LDS contents do not matter.
template <class Probe = NoProbe, bool kC = false>
__device__ inline uint32_t dequeue(Ring* r, Probe* pr = nullptr, bool idle_check = true, Combine* cb = nullptr) {
  const Geometry g = geom<kC>(r);
  uint64_t* const e = slots(r);
  __shared__ uint8_t s_low[16];  // per wave: the free ring was found empty lately
  const uint32_t wv = ((threadIdx.x + threadIdx.y * blockDim.x) / kWave) & 15u;
  bool bounded = idle_check;
  if (!idle_check) {
    if (ld_rlx(thr_p<kC>(r)) < 0) {  // consumer side: the claim tests emptiness
      s_low[wv] = low_next(s_low[wv], true);
      return kBot;
    }
    bounded = (s_low[wv] & 15u) != 0;
  }
  if (idle_check) idle_pause(g.poll_backoff);
  for (;;) {
    uint64_t H;
    if (bounded) {
      // Combined with the workgroup's other idle waves when cb is given.
      if (cb) {
        uint8_t* const recent = combine_recent();
        const uint32_t rc = *recent & 7u;
        H = rc ? claim_head_combined<kC>(r, cb) : claim_head<kC>(r);
        *recent = (uint8_t)(H == kEmptyTicket ? kCombineAfterEmpty : (rc ? rc - 1u : 0u));
      } else {
        H = claim_head<kC>(r);
      }
      if (H == kEmptyTicket) {
        if (idle_check) idle_note(g.poll_backoff, true);
        else s_low[wv] = low_next(s_low[wv], true);
        return kBot;
      }
      if (H == kNoTicket) {
        if (pr) pr->step();
        continue;
      }
    } else {
      H = wave_faa(head_p<kC>(r), (uint64_t)1);
    }
    if (pr) { pr->step(); pr->after_ticket(); }
    const uint32_t hcyc = cycle_of(H, g.slots_log2);
    const uint32_t j = slot_of(H, g.slots_log2, g.remap_log2);
    uint64_t E = ld_rlx(&e[j]);
    // A bounded ticket lies below Tail: its enqueuer is in flight; give it a
    // moment before marking the slot.
    for (uint32_t w = 0; bounded && w < kSlotWaitSpins && cyc_behind(e_cyc(E), hcyc) && e_idx(E) >= kBot; ++w) {
      __builtin_amdgcn_s_sleep(1);
      E = ld_rlx(&e[j]);
      if (pr) pr->step();
    }
    for (;;) {
      const uint32_t ecyc = e_cyc(E), eidx = e_idx(E);
      if (ecyc == hcyc && eidx < kBot) {
        or_rlx(&e[j], (uint64_t)kBotC);  // consume: index := ⊥c, keep cycle/safe
        if (idle_check) idle_note(g.poll_backoff, false);
        else s_low[wv] = low_next(s_low[wv], false);
        return eidx;
      }
      const uint64_t nv = (eidx >= kBot) ? e_pack(hcyc, e_safe(E), kBot) : e_pack(ecyc, false, eidx);
      if (cyc_behind(ecyc, hcyc) && !cas_rlx(&e[j], E, nv)) {
        if (pr) pr->step();
        continue;
      }
      break;
    }
    const uint64_t T = ld_rlx(tail_p<kC>(r));
    if (T <= H + 1) {
      catchup<kC>(r, H + 1);
      faa_rlx(thr_p<kC>(r), -1);
      if (!idle_check) s_low[wv] = low_next(s_low[wv], true);
      return kBot;
    }
    if (faa_rlx(thr_p<kC>(r), -1) <= 0) {
      if (!idle_check) s_low[wv] = low_next(s_low[wv], true);
      return kBot;
    }
  }
}

// ---------------------------------------------------------------- mpmc
// u64 payload queue: aq holds published indices, fq free indices, data the
// payloads (reference fq_mpmc_t, glf_queue.hpp:134-138).
struct Mpmc {
  Ring* aq;
  Ring* fq;
  uint64_t* data;
};
__host__ __device__ inline uint64_t mpmc_bytes(uint32_t n) { return 2 * ring_bytes(n) + 8ull * n; }
__host__ __device__ inline Mpmc mpmc_view(void* base, uint32_t n) {
  Mpmc m;
  m.aq = (Ring*)base;
  m.fq = (Ring*)((char*)base + ring_bytes(n));
  m.data = (uint64_t*)((char*)base + 2 * ring_bytes(n));
  return m;
}

template <class Probe = NoProbe>
__device__ inline bool enqueue_ptr(const Mpmc& q, uint64_t v, Probe* pr = nullptr, Combine* ct = nullptr) {
  const uint32_t idx = dequeue(q.fq, pr, false);
  if (idx == kBot) return false;  // full
  st_rlx(&q.data[idx], v);  // acknowledged before enqueue() touches a slot
  enqueue(q.aq, idx, pr, false, ct);
  return true;
}

template <class Probe = NoProbe>
__device__ inline bool dequeue_ptr(const Mpmc& q, uint64_t* out, Probe* pr = nullptr, Combine* cb = nullptr) {
  const uint32_t idx = dequeue(q.aq, pr, true, cb);
  if (idx == kBot) return false;  // empty
  *out = ld_rlx(&q.data[idx]);  // address-dependent on the dequeued index;
  enqueue(q.fq, idx, pr, true);  // completes before enqueue() touches a slot
  return true;
}

}  // namespace lfq
}  // namespace gcq
