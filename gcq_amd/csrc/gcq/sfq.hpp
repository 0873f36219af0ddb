// sfq.hpp — SFQ: Scogland & Feng ticket ring (ICPE'15), blocking and
// non-blocking forms, wave64-native.
//
// Algorithm follows the reference sfqueue_hip.{hpp,cpp}:
//   ticket t -> slot t % QLEN, phase id = 2*(t / QLEN) (+1 for dequeue)
//   blocking ops  FAA ticket, wait for ids[slot] == id     sfqueue_hip.cpp:55-87
//   non-blocking  load ticket, check id, CAS ticket        sfqueue_hip.cpp:89-128
//   closed flag unblocks every wait                        sfqueue_hip.cpp:47-53
// Own MI355X design choices:
//  * 64-bit tickets, phase ids and items.  The reference's 32-bit invariant
//    (sfqueue_hip.hpp:45-47) caps the thread count at 4,095 for QLEN = 2^20;
//    64-bit phases never wrap, so 65,536 threads run at any capacity.
//  * slot = {id, item} in one 16-byte pair, so the phase check and the item
//    transfer touch one line.
//  * blocking tickets come from one aggregated FAA per wave; non-blocking
//    claims are wave-batched: the active lanes test consecutive tickets
//    t0+rank and one CAS per wave claims the ready prefix [t0, t0+k), or
//    one CAS per workgroup when the caller passes a Combine (nb_claim_group).
//  * Tail, Head and the closed flag sit on separate 128-B lines.
#pragma once
#include "prim.hpp"

namespace gcq {
namespace sfq {

constexpr int kSuccess = 0, kBusy = 1, kClosed = 2;  // sfqueue_hip.hpp:28-30
constexpr uint32_t kHdrBytes = 2048;

struct Slot {
  uint64_t id;
  uint64_t item;
};

struct Geometry {
  uint64_t qlen;
  uint32_t qlen_log2, nprocs;
  uint32_t off_tail, off_head, off_closed, magic;
  uint32_t backoff;  // polite-wait iterations between checks (ref SFQ_BACKOFF_ITERS)
  uint32_t pad;
};
constexpr uint32_t kMagic = 0x53465131u;  // "SFQ1"
struct Queue {
  Geometry g;
  uint8_t pad_[kHdrBytes - sizeof(Geometry)];
};
static_assert(sizeof(Queue) == kHdrBytes, "header size");

__host__ __device__ constexpr uint64_t queue_bytes(uint64_t qlen) { return kHdrBytes + 16ull * qlen; }

__device__ __forceinline__ Slot* slots(Queue* q) { return (Slot*)((char*)q + kHdrBytes); }
__device__ __forceinline__ uint64_t* tail(Queue* q) { return (uint64_t*)((char*)q + ld_const(&q->g).off_tail); }
__device__ __forceinline__ uint64_t* head(Queue* q) { return (uint64_t*)((char*)q + ld_const(&q->g).off_head); }
__device__ __forceinline__ uint32_t* closed(Queue* q) { return (uint32_t*)((char*)q + ld_const(&q->g).off_closed); }

struct NoProbe {
  __device__ __forceinline__ void step() {}
  __device__ __forceinline__ void after_ticket() {}
  __device__ __forceinline__ void slow_entry() {}
};

__device__ inline void init_slice(Queue* q, uint64_t qlen, uint32_t nprocs, uint32_t gid, uint32_t stride) {
  Slot* s = slots(q);
  for (uint64_t i = gid; i < qlen; i += stride) {
    s[i].id = 0;
    s[i].item = 0;
  }
  if (gid == 0) {
    Geometry g;
    g.qlen = qlen;
    g.qlen_log2 = 63u - (uint32_t)__builtin_clzll(qlen);
    g.nprocs = nprocs;
    g.off_tail = 1 * kLine;
    g.off_head = 5 * kLine;
    g.off_closed = 9 * kLine;
    g.magic = kMagic;
    g.backoff = 64;
    g.pad = 0;
    q->g = g;
    *(uint64_t*)((char*)q + g.off_tail) = 0;
    *(uint64_t*)((char*)q + g.off_head) = 0;
    *(uint32_t*)((char*)q + g.off_closed) = 0;
  }
}

__device__ __forceinline__ bool is_closed(Queue* q) { return ld_rlx(closed(q)) != 0u; }
__device__ __forceinline__ void close(Queue* q) { st_rlx(closed(q), 1u); }

__device__ __forceinline__ void backoff(uint32_t iters) {
  for (uint32_t i = 0; i < iters; ++i) __builtin_amdgcn_s_sleep(1);
}
// Wait between checks of a slot's phase in the blocking operations: short
// for the first kQuickChecks checks (the slot's other party is usually
// already past its ticket), then the queue's polite backoff.
constexpr uint64_t kQuickChecks = 64;
__device__ __forceinline__ void slot_wait(uint64_t spins, uint32_t iters) {
  if (spins < kQuickChecks) __builtin_amdgcn_s_sleep(1);
  else backoff(iters);
}

// Blocking enqueue (sfqueue_hip.cpp:55-70).  `spin_limit` bounds the wait
// (0 = unbounded); exceeding it returns kBusy so no caller hangs forever.
template <class Probe = NoProbe>
__device__ inline int enqueue_blocking(Queue* q, uint64_t item, uint64_t spin_limit = 0, Probe* pr = nullptr) {
  if (is_closed(q)) return kClosed;
  const Geometry g = ld_const(&q->g);
  const uint64_t t = wave_faa(tail(q), (uint64_t)1);
  if (pr) { pr->step(); pr->after_ticket(); }
  Slot* s = &slots(q)[t & (g.qlen - 1)];
  const uint64_t id = 2ull * (t >> g.qlen_log2);
  uint64_t spins = 0;
  while (ld_rlx(&s->id) != id) {  // the previous lap's dequeuer has released the slot
    if (is_closed(q)) return kClosed;
    if (spin_limit && ++spins > spin_limit) return kBusy;
    slot_wait(spins, g.backoff);
  }
  st_rlx(&s->item, item);
  vm_wait();  // item acknowledged before the phase flip publishes it
  st_rlx(&s->id, id + 1);
  return kSuccess;
}

template <class Probe = NoProbe>
__device__ inline int dequeue_blocking(Queue* q, uint64_t* out, uint64_t spin_limit = 0, Probe* pr = nullptr) {
  if (is_closed(q)) return kClosed;
  const Geometry g = ld_const(&q->g);
  const uint64_t t = wave_faa(head(q), (uint64_t)1);
  if (pr) { pr->step(); pr->after_ticket(); }
  Slot* s = &slots(q)[t & (g.qlen - 1)];
  const uint64_t id = 2ull * (t >> g.qlen_log2) + 1ull;
  uint64_t spins = 0;
  while (ld_rlx(&s->id) != id) {
    if (is_closed(q)) return kClosed;
    if (spin_limit && ++spins > spin_limit) return kBusy;
    slot_wait(spins, g.backoff);
  }
  *out = ld_rlx(&s->item);
  vm_wait();  // item read completes before the slot is handed to the next lap
  st_rlx(&s->id, id + 1);
  return kSuccess;
}

// Wave-batched non-blocking claim on counter *ctr: lanes test consecutive
// tickets; one CAS claims the ready prefix.  Returns this lane's ticket or
// ~0 when its op is BUSY.
// Contention backoff of the non-blocking claim, per wave: after a lost CAS
// on the ticket counter the wave pauses 2^e x 64 cycles before answering
// BUSY, with e raised by one per loss (up to kBusyBackoffMax) and lowered by
// one per won claim.  Every answer is still one attempt, so the operations
// stay non-blocking; the pause only thins out the waves that contend for the
// counter line at once.  Measured, split 25/75 at 1,024 / 8,192 / 65,536
// threads: 1.2x / 1.9x / 5.5x the successful operations per second without
// it.  The exponent lives in LDS and is clamped on every read, so its
// contents at launch only bound the first pause.
// The cap follows the number of contenders: log2 of the grid's waves, less
// one, within [3, kBusyBackoffMax].  Measured with fixed caps, split 25/75:
// cap 3 is best at 1,024 threads (152 against 112 Mops/s with 9), 7 at
// 8,192 (134 against 129), 9 from 32,768 up.  The scaled cap against a fixed
// 9, alternating runs: 1,024 / 2,048 / 4,096 threads +17% / +19% / +15%,
// level from 8,192 up.
constexpr uint32_t kBusyBackoffMax = 9;
__device__ __forceinline__ uint32_t busy_backoff_cap() {
  const uint32_t waves = gridDim.x * gridDim.y * gridDim.z * blockDim.x * blockDim.y * blockDim.z / kWave;
  const uint32_t l = 31u - (uint32_t)__builtin_clz(waves | 1u);
  return l < 4u ? 3u : (l - 1u < kBusyBackoffMax ? l - 1u : kBusyBackoffMax);
}
__device__ __forceinline__ uint8_t* busy_backoff() {
  __shared__ uint8_t s_bo[16];
  return &s_bo[((threadIdx.x + threadIdx.y * blockDim.x) / kWave) & 15u];
}

__device__ __forceinline__ uint64_t nb_claim(Queue* q, uint64_t* ctr, uint64_t phase_add) {
  const Geometry g = ld_const(&q->g);
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  uint64_t t0 = 0;
  if (lane_id() == leader) t0 = ld_rlx(ctr);
  t0 = bcast64(t0, leader);
  const uint64_t t = t0 + rank_in(m);
  const uint64_t id = 2ull * (t >> g.qlen_log2) + phase_add;
  const bool ready = ld_rlx(&slots(q)[t & (g.qlen - 1)].id) == id;
  const uint64_t not_ready = m & ~__ballot(ready);
  const uint32_t k = not_ready ? __builtin_popcountll(m & ((1ull << first_lane(not_ready)) - 1ull))
                               : (uint32_t)__builtin_popcountll(m);
  uint32_t won = 0;
  if (lane_id() == leader && k > 0) {
    uint64_t exp = t0;
    won = cas_rlx(ctr, exp, t0 + k) ? 1u : 0u;
  }
  won = bcast32(won, leader);
  if (k > 0) {
    uint8_t* const bo = busy_backoff();
    const uint32_t cap = busy_backoff_cap(), e0 = *bo & 15u, e = e0 < cap ? e0 : cap;
    if (!won) {
      for (uint32_t i = 0; i < (1u << e); ++i) __builtin_amdgcn_s_sleep(1);
      if (lane_id() == leader) *bo = (uint8_t)(e < cap ? e + 1u : e);
    } else if (lane_id() == leader) {
      *bo = (uint8_t)(e ? e - 1u : 0u);
    }
  }
  return (won && rank_in(m) < k) ? t : ~0ull;
}

// The non-blocking claim shared by the waves of a workgroup that claim at
// about the same time (the same LDS group protocol as combine_claim in
// prim.hpp: join, the leader closes the group after the previous group's
// claim, members wait for the leader's result).  The leading wave reads the
// counter once, tests the phase of the group's N tickets t0 .. t0 + N - 1
// with all its active lanes (up to 4 loads per lane, issued together), and
// claims the ready prefix [t0, t0 + k) with one CAS; lane r of a member at
// offset o gets t0 + o + r when o + r < k.  One CAS per workgroup instead of
// one per wave thins the contention on the counter line, and a won CAS hands
// out up to a whole workgroup's tickets.  Every member's operation is in
// progress between the read and the CAS, as in the per-wave claim.
constexpr uint32_t kGroupProbeRounds = 4;
__device__ inline uint64_t nb_claim_group(Queue* q, uint64_t* ctr, uint64_t phase_add, Combine* c) {
  const Geometry g = ld_const(&q->g);
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  const uint32_t n = (uint32_t)__builtin_popcountll(m);
  const uint32_t rank = rank_in(m);
  const uint32_t w = wave_in_block();
  uint32_t grp = 0, off = 0;
  if (lane_id() == leader) {
    __hip_atomic_store(&c->res_g[w], ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint64_t old = __hip_atomic_fetch_add(&c->join, (1ull << (32 + w)) + n, __ATOMIC_RELEASE,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
    grp = (uint32_t)(old >> 48);
    off = (uint32_t)old;
  }
  grp = bcast32(grp, leader);
  off = bcast32(off, leader);
  uint32_t k = 0;
  uint64_t base = 0;
  if (off == 0) {
    uint64_t fin = 0;
    if (lane_id() == leader) {
      for (uint32_t sp = 0; sp < kCombineLeaderSpins &&
                            (__hip_atomic_load(&c->done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) & 0xFFFFu) !=
                                ((grp - 1u) & 0xFFFFu);
           ++sp)
        __builtin_amdgcn_s_sleep(1);
      fin = __hip_atomic_exchange(&c->join, (uint64_t)((grp + 1u) & 0xFFFFu) << 48, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
      lgkm_wait();  // the group is closed before the counter is read
      base = ld_rlx(ctr);
    }
    fin = bcast64(fin, leader);
    base = bcast64(base, leader);
    const uint32_t total = (uint32_t)fin;
    const uint32_t cover = total < kGroupProbeRounds * n ? total : kGroupProbeRounds * n;
    bool ready[kGroupProbeRounds];
#pragma unroll
    for (uint32_t r = 0; r < kGroupProbeRounds; ++r) {
      const uint32_t j = r * n + rank;
      const uint64_t t = base + j;
      ready[r] = j >= cover || ld_rlx(&slots(q)[t & (g.qlen - 1)].id) == 2ull * (t >> g.qlen_log2) + phase_add;
    }
    k = cover;
#pragma unroll
    for (uint32_t r = 0; r < kGroupProbeRounds; ++r) {
      const uint64_t bad = m & ~__ballot(ready[r]);
      if (bad && k == cover) {  // the smallest unready ticket of the round: the lowest unready lane's rank
        const uint32_t fl = first_lane(bad);
        k = r * n + (uint32_t)__builtin_popcountll(m & ((1ull << fl) - 1ull));
      }
    }
    uint32_t won = 0;
    if (lane_id() == leader && k > 0) {
      uint64_t exp = base;
      won = cas_rlx(ctr, exp, base + k) ? 1u : 0u;
    }
    won = bcast32(won, leader);
    if (k > 0) {  // the per-wave contention backoff, for the leading wave
      uint8_t* const bo = busy_backoff();
      const uint32_t cap = busy_backoff_cap(), e0 = *bo & 15u, e = e0 < cap ? e0 : cap;
      if (!won) {
        for (uint32_t i = 0; i < (1u << e); ++i) __builtin_amdgcn_s_sleep(1);
        if (lane_id() == leader) *bo = (uint8_t)(e < cap ? e + 1u : e);
      } else if (lane_id() == leader) {
        *bo = (uint8_t)(e ? e - 1u : 0u);
      }
    }
    if (!won) k = 0;
    if (lane_id() == leader) {
      for (uint32_t b = (uint32_t)(fin >> 32) & 0xFFFFu & ~(1u << w); b; b &= b - 1) {
        const uint32_t x = (uint32_t)__builtin_ctz(b);
        c->res_k[x] = k;
        c->res_base[x] = base;
        __hip_atomic_store(&c->res_g[x], grp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      __hip_atomic_store(&c->done, grp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  } else {
    if (lane_id() == leader) {
      while (__hip_atomic_load(&c->res_g[w], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != grp)
        __builtin_amdgcn_s_sleep(1);
      k = c->res_k[w];
      base = c->res_base[w];
    }
    k = bcast32(k, leader);
    base = bcast64(base, leader);
  }
  return off + rank < k ? base + off + rank : ~0ull;
}

// Non-blocking enqueue (sfqueue_hip.cpp:89-107).  With a Combine the claim
// is shared by the workgroup's enqueuing waves (nb_claim_group).
template <class Probe = NoProbe>
__device__ inline int enqueue_nb(Queue* q, uint64_t item, Probe* pr = nullptr, Combine* cg = nullptr) {
  if (is_closed(q)) return kClosed;
  const uint64_t t = cg ? nb_claim_group(q, tail(q), 0, cg) : nb_claim(q, tail(q), 0);
  if (pr) pr->step();
  if (t == ~0ull) return kBusy;
  Slot* s = &slots(q)[t & (ld_const(&q->g).qlen - 1)];
  st_rlx(&s->item, item);
  vm_wait();
  st_rlx(&s->id, 2ull * (t >> ld_const(&q->g).qlen_log2) + 1ull);
  return kSuccess;
}

// Non-blocking dequeue (sfqueue_hip.cpp:109-128).  With a Combine the claim
// is shared by the workgroup's dequeuing waves.
template <class Probe = NoProbe>
__device__ inline int dequeue_nb(Queue* q, uint64_t* out, Probe* pr = nullptr, Combine* cg = nullptr) {
  if (is_closed(q)) return kClosed;
  const uint64_t t = cg ? nb_claim_group(q, head(q), 1, cg) : nb_claim(q, head(q), 1);
  if (pr) pr->step();
  if (t == ~0ull) return kBusy;
  Slot* s = &slots(q)[t & (ld_const(&q->g).qlen - 1)];
  *out = ld_rlx(&s->item);
  vm_wait();
  st_rlx(&s->id, 2ull * (t >> ld_const(&q->g).qlen_log2) + 2ull);
  return kSuccess;
}

// Dequeue that reports EMPTY only when it has observed head >= tail (the
// queue was empty at the tail read); BUSY attempts are retried.  This is the
// BFS adapter rule of the reference (bfs_bench.cpp:114-131) and gives
// dequeue results the FIFO model can check.
template <class Probe = NoProbe>
__device__ inline bool dequeue_or_empty(Queue* q, uint64_t* out, uint64_t retry_limit, Probe* pr = nullptr,
                                        Combine* cg = nullptr) {
  for (uint64_t k = 0;; ++k) {
    const int rc = dequeue_nb(q, out, pr, cg);
    if (rc == kSuccess) return true;
    if (rc == kClosed) return false;
    const uint64_t h = ld_rlx(head(q));
    vm_wait();  // head is read before tail: head(tail read) >= h >= t means empty
    const uint64_t t = ld_rlx(tail(q));
    if (h >= t) return false;
    if (retry_limit && k >= retry_limit) return false;
    backoff(8);
  }
}

}  // namespace sfq
}  // namespace gcq
