// prim.hpp — memory-order and wave64 primitives shared by every gcq queue.
//
// Every shared access in the queues goes through one of these helpers so that
// each site states its order explicitly.  All atomics are agent scope (one
// GPU, all 8 XCDs of an MI355X).  As lowered by hipcc for gfx950:
//   relaxed RMW      -> global_atomic_* (performed at the coherent L2/EA point)
//   relaxed load     -> global_load ... sc1          (coherent read)
//   relaxed store    -> global_store ... sc1         (write-through)
//   release fence    -> buffer_wbl2 sc1 + s_waitcnt  (publish dirty L2 lines)
//   acquire fence    -> buffer_inv sc1               (drop stale lines)
// The reference relies on HIP atomic* being seq_cst (gwf_ring.hpp:15-17);
// they are relaxed, so every cross-thread hand-off here states its order: a
// completion wait (vm_wait, below) on hot paths, fences on slow paths only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gcq {

constexpr uint32_t kWave = 64;        // CDNA wavefront width
constexpr uint32_t kLine = 128;       // MI355X L2 line size in bytes

// ---------------------------------------------------------------- atomics
// Value parameters are non-deduced so that literal operands convert to the
// pointee type.
template <class T>
struct nd { using type = T; };
template <class T>
using nd_t = typename nd<T>::type;

template <class T>
__device__ __forceinline__ T ld_rlx(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Relaxed load that may be served by this CU's or XCD's caches (workgroup
// scope): for words that change at most once during a kernel and only one
// way (a depth going from -1 to a level), where a stale value costs no more
// than a redundant atomic, and for words that do not change at all during
// the kernel.
template <class T>
__device__ __forceinline__ T ld_cached(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Read of a queue's geometry header: written only by the queue's init
// kernels, never while another kernel uses the queue, so it is read through
// the constant address space.  The compiler may then load it with scalar
// loads and keep it across a kernel's loop; a plain pointer made it reload
// the header with a vector load and wait before every operation's first
// atomic, since any store or atomic in the loop might have changed it.
// Init code, which writes the header and reads it back in the same kernel,
// reads it plainly.  Used by SFQ, WFQ and the BFS frontier adapters.  The
// G-WFQ and G-LFQ rings keep plain reads: measured in the split 25/75
// throughput run, constant reads there gained 7-16% at 16,384 and 32,768
// threads but lost 15-17% at 65,536 with claims combined per workgroup
// (3,230 -> 2,600 Mops/s for G-WFQ), and neither a sleep in their place nor
// consumer backoff recovered it.
template <class T>
__device__ __forceinline__ T ld_const(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  typedef const __attribute__((address_space(4))) uint32_t* W;
  const W w = (W)p;
  T out;
  uint32_t* o = (uint32_t*)&out;
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i) o[i] = w[i];  // unused words are dropped by the compiler
  return out;
}
// Agent-scope acquire after a successful dequeue through the drop-in
// headers' index and value APIs, for callers that read the dequeued index's
// payload with plain loads (the ray tracing harness publishes ray records
// with plain stores and __threadfence and reads them with plain loads after
// the dequeue).  Every queue operation here is relaxed, so without it such a
// load may be served by a cache line this CU or XCD fetched before the
// record was rewritten: measured, 1 to 4 of 1,712,002 rays per frame were
// traced again from a stale record.  The acquire costs about 10% of that
// harness's trace time; GCQ_DROPIN_RELAXED leaves it out.
__device__ __forceinline__ void dropin_acquire() {
#ifndef GCQ_DROPIN_RELAXED
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#endif
}
// System scope: words in host-pinned memory shared with host threads.
template <class T>
__device__ __forceinline__ T ld_sys(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ __forceinline__ void st_sys(T* p, nd_t<T> v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ __forceinline__ T ld_acq(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rlx(T* p, nd_t<T> v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rel(T* p, nd_t<T> v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T faa_rlx(T* p, nd_t<T> v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T or_rlx(T* p, nd_t<T> v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T max_rlx(T* p, nd_t<T> v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T xchg_rlx(T* p, nd_t<T> v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// CAS that refreshes `expected` with the observed value on failure.
template <class T>
__device__ __forceinline__ bool cas_rlx(T* p, T& expected, nd_t<T> desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// CAS with release on success: everything stored before it is visible to an
// agent-scope acquirer that observes the new value.
template <class T>
__device__ __forceinline__ bool cas_rel(T* p, T& expected, nd_t<T> desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELEASE,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ bool cas_once(T* p, nd_t<T> expected, nd_t<T> desired) {
  return cas_rlx(p, expected, desired);
}

// Full agent-scope fences.  Measured on MI355X at 65,536 threads: a release
// or acquire fence costs ~7.4 us per wave (the L2 writeback / invalidate is
// serialised device-wide) against ~0.08 us for a relaxed store, so they are
// kept off every hot path (slow-path request publication only).
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_sc() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent"); }

// Completion wait for this lane's outstanding vector-memory operations.
// Hot-path payload hand-off uses it instead of a fence: payload words are
// written and read only with agent-scope atomic stores/loads (sc1: performed
// at the device-coherent level, never left dirty in a per-XCD L2), so once
// the store is acknowledged it is visible to every XCD, and a consumer's
// payload load is address-dependent on the index it dequeued.  Waiting for
// completion therefore orders "payload store -> index publish" and "payload
// load -> index recycle" without the L2 writeback/invalidate of a fence.
// The "memory" clobber also stops the compiler moving accesses across it.
__device__ __forceinline__ void vm_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ---------------------------------------------------------------- wave64
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t active_mask() { return __ballot(1); }
__device__ __forceinline__ uint64_t lanes_below() {
  // __builtin_amdgcn_mbcnt_{lo,hi} counts set bits of a mask below this lane.
  return (uint64_t)__builtin_amdgcn_mbcnt_hi(0xffffffffu, __builtin_amdgcn_mbcnt_lo(0xffffffffu, 0u));
}
__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t first_lane(uint64_t mask) { return (uint32_t)__builtin_ctzll(mask); }
__device__ __forceinline__ uint64_t bcast64(uint64_t v, uint32_t src) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t bcast32(uint32_t v, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}

// One device-scope fetch-and-add per wave: the active lanes elect the lowest
// lane, which adds (popcount * step); each lane then receives
// old + rank * step where rank is its position among the active lanes.
// Same contract as the reference's aggregated ticket FAA
// (gwf_ring.hpp:202-223, glf_queue.hpp:392-412).
template <class T>
__device__ __forceinline__ T wave_faa(T* p, nd_t<T> step) {
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  const uint32_t rank = rank_in(m);
  T old = 0;
  if (lane_id() == leader) old = faa_rlx(p, (T)(__builtin_popcountll(m)) * step);
  old = (T)bcast64((uint64_t)old, leader);
  return old + (T)rank * step;
}

// ------------------------------------------------------ workgroup combining
// The consumer-side claim reads Head and then Tail, two lines that every
// producer and consumer wave of the grid also hits.  Measured on MI355X, one
// line serves about 83 M device-scope FAAs/s; once loads of it are mixed with
// FAAs, each load costs about as much as an FAA, and the line tops out at
// 85-100 M accesses/s in all.  The polling of idle consumers therefore delays
// the producers' Tail FAAs.  Combining lets the waves of one workgroup that
// claim at about the same time share one Head/Tail read pair and one FAA.
// The protocol runs entirely in LDS:
//  - a wave joins the open group with one 64-bit LDS add.  The add sets the
//    wave's member bit and adds its lane count, and returns the group id and
//    the lanes ahead of the wave (its offset);
//  - the first wave of a group leads it.  It waits until the previous
//    group's claim has finished, so that later waves can join meanwhile.
//    Then it closes the group by swapping in the next group id, and only
//    then reads Head and Tail;
//  - every member joined before the close, and so before the reads.  An EMPTY
//    answer (Tail <= Head) is therefore a witness inside every member's
//    operation;
//  - the leader takes k = min(lanes, Tail - Head) tickets with one FAA and
//    hands (k, base) to each member's LDS slot.  Lane r of a member at offset
//    o gets ticket base + o + r when o + r < k.
// All waits are on LDS, between waves of one workgroup, which are all
// resident.  The leader of group g waits (boundedly) only for the leader of
// group g - 1, which has already closed g - 1; a member waits only for its
// group's leader.  A kernel that
// passes a Combine to a dequeue calls combine_init() from one thread, then
// a barrier, before any dequeue.
struct Combine {
  uint64_t join;       // [63:48] open group id, [47:32] member waves, [31:0] lanes
  uint32_t done;       // id of the last group whose claim has finished
  uint32_t res_g[16];  // per wave: group whose result res_k/res_base hold
  uint32_t res_k[16];
  uint64_t res_base[16];
};
__device__ __forceinline__ void combine_init(Combine* c) {
  c->join = 1ull << 48;
  c->done = 0;
  for (int i = 0; i < 16; ++i) c->res_g[i] = ~0u;
}
__device__ __forceinline__ uint32_t wave_in_block() {
  return ((threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y) / kWave) & 15u;
}
__device__ __forceinline__ void lgkm_wait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// A wave combines only while it keeps finding the queue empty: for
// kCombineAfterEmpty claims after its last EMPTY answer.  Idle consumers are
// the waves whose polling loads the Head/Tail lines.  Waves that find items
// claim alone, so a busy queue does not wait on group turns.  The per-wave
// count is only a choice of path; its LDS contents at launch do not matter.
constexpr uint32_t kCombineAfterEmpty = 4;
// Smallest grid worth combining for.  With fewer polling waves the hot lines
// are not congested and a group's turn only adds latency.  Measured, split
// 25/75, block 256, combined over solo: G-WFQ 0.87 / 1.00 / 1.03 / 1.04 /
// 1.08 at 2,048 / 4,096 / 8,192 / 16,384 / 65,536 threads; G-LFQ 0.97 /
// 0.99 / 1.03 / 1.05 / 1.05.
constexpr uint32_t kCombineMinThreads = 8192;
__device__ __forceinline__ uint8_t* combine_recent() {
  __shared__ uint8_t s_recent[16];
  return &s_recent[wave_in_block()];
}

constexpr uint32_t kCombineLeaderSpins = 1u << 20;

// Run by the calling wave's leader lane only.  `memclaim(lanes, &k, &base)`
// runs in the group leader after the close: it reads Head, then Tail, and
// takes k tickets starting at base.  Returns this wave's offset in the group;
// *k and *base are the group's.
template <class F>
__device__ __forceinline__ uint32_t combine_claim(Combine* c, uint32_t lanes, uint32_t* k, uint64_t* base,
                                                  F&& memclaim) {
  const uint32_t w = wave_in_block();
  // Group ids are 16 bits wide.  A wave that was last a member of group X
  // and next joins group X + 65,536 (having led or sat out every group in
  // between) would otherwise find res_g[w] == X already and take that old
  // group's result.  Only the leader of a group this wave joined writes
  // res_g[w], and that leader has already written it for the wave's last
  // membership, so the clear cannot race; the release join orders it before
  // the leader's close.
  __hip_atomic_store(&c->res_g[w], ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint64_t old = __hip_atomic_fetch_add(&c->join, (1ull << (32 + w)) + lanes, __ATOMIC_RELEASE,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint32_t grp = (uint32_t)(old >> 48);
  const uint32_t off = (uint32_t)old;
  if (off == 0) {
    // The wait only lets more waves join; the answer's validity rests on the
    // close below.  Bounding it means every leader reaches its publish, so no
    // member waits on a leader that is itself waiting.
    for (uint32_t s = 0; s < kCombineLeaderSpins &&
                         (__hip_atomic_load(&c->done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) & 0xFFFFu) !=
                             ((grp - 1u) & 0xFFFFu);
         ++s)
      __builtin_amdgcn_s_sleep(1);
    const uint64_t fin = __hip_atomic_exchange(&c->join, (uint64_t)((grp + 1u) & 0xFFFFu) << 48, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
    lgkm_wait();  // the group is closed before Head is read
    memclaim((uint32_t)fin, k, base);
    for (uint32_t b = (uint32_t)(fin >> 32) & 0xFFFFu & ~(1u << w); b; b &= b - 1) {
      const uint32_t x = (uint32_t)__builtin_ctz(b);
      c->res_k[x] = *k;
      c->res_base[x] = *base;
      __hip_atomic_store(&c->res_g[x], grp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __hip_atomic_store(&c->done, grp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    while (__hip_atomic_load(&c->res_g[w], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != grp)
      __builtin_amdgcn_s_sleep(1);
    *k = c->res_k[w];
    *base = c->res_base[w];
  }
  return off;
}

// Producers' ticket FAA on the published ring's Tail, shared by the waves of
// a workgroup that draw together: the leader adds the group's total lane
// count once and each lane gets old + (its offset in the group) * step.  The
// tickets are the ones a single wave of that many lanes would draw, taken
// while every member's operation is in progress.  The Tail line is the one
// idle consumers poll, so fewer FAAs there shorten every wave's wait on it.
// Measured, G-WFQ, block 256, grouped over per-wave FAAs: split 25/75 at
// 8,192 / 16,384 / 65,536 threads 0.96 / 0.99 / 1.11, balanced at 65,536
// 1.04.  Grouping the free-index ring's Head FAA as well (only producers
// use that line) ran 16% slower at 65,536 than grouping the Tail FAA alone.
constexpr uint32_t kCombineEnqMinThreads = 32768;
template <class T>
__device__ __forceinline__ T group_faa(Combine* c, T* p, nd_t<T> step) {
  const uint64_t m = active_mask();
  const uint32_t leader = first_lane(m);
  const uint32_t rank = rank_in(m);
  const uint32_t n = (uint32_t)__builtin_popcountll(m);
  uint64_t base = 0;
  uint32_t off = 0;
  if (lane_id() == leader) {
    uint32_t k = 0;
    off = combine_claim(c, n, &k, &base, [&](uint32_t lanes, uint32_t* kk, uint64_t* bb) {
      *kk = lanes;
      *bb = (uint64_t)faa_rlx(p, (T)lanes * step);
    });
  }
  off = bcast32(off, leader);
  base = bcast64(base, leader);
  return (T)base + (T)(off + rank) * step;
}

// ------------------------------------------------- free-ring threshold reset
// SCQ resets a ring's threshold to 3n - 1 after every insertion: "if it is
// not 3n - 1, store 3n - 1" (a load, then a store only if needed).  The
// published ring's enqueuers are the producers, on the throughput path, so
// they store unconditionally, once per wave.  The free-index ring's enqueuers
// are consumers returning indices.  They use SCQ's form, so the threshold
// line stays quiet while no producer fails on that ring; every producer
// reads it before drawing a free index, and a load of a line that is also
// being written costs about as much as an atomic there.  Measured, split
// 25/75 at 65,536 threads: G-WFQ +15%.  A wave that found the threshold
// below 3n - 1 stores unconditionally for its next kResetStreak returns,
// since the free ring is running low then.  The per-wave count is only a
// choice of form (both are SCQ resets), so its LDS contents at launch do not
// matter.
constexpr uint32_t kResetStreak = 8;
__device__ __forceinline__ uint8_t* reset_streak() {
  __shared__ uint8_t s_hot[16];
  return &s_hot[((threadIdx.x + threadIdx.y * blockDim.x) / kWave) & 15u];
}
// One threshold reset after a wave's insertions, by one lane.
__device__ __forceinline__ void reset_threshold(int32_t* thr, int32_t max, bool scq_form) {
  if (!scq_form) {
    st_rlx(thr, max);
    return;
  }
  uint8_t* const hot = reset_streak();
  const uint32_t h = *hot & 15u;
  if (h) {
    st_rlx(thr, max);
    *hot = (uint8_t)(h - 1u);
  } else if (ld_rlx(thr) != max) {
    st_rlx(thr, max);
    *hot = (uint8_t)kResetStreak;
  }
}

// Free-index ring hint: after a wave finds the free ring empty, its next
// kLowStreak free-ring dequeues take bounded tickets (claims) instead of
// plain aggregated ones.  Plain tickets on an empty or nearly empty free
// ring mark slots ahead of the consumers returning indices.  Holding the
// bounded form for a few operations, instead of dropping it at the first
// success, doubles split 75/25 throughput on a full 65,536-item ring.
constexpr uint32_t kLowStreak = 8;
__device__ __forceinline__ uint8_t low_next(uint8_t cur, bool found_empty) {
  const uint32_t c = cur & 15u;
  return (uint8_t)(found_empty ? kLowStreak : (c ? c - 1u : 0u));
}

// Wave-wide sum of a 64-bit value; all 64 lanes must be active (the
// shuffles read every lane).  The result is valid in every lane.
__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)v, off, 64);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Bounded wait, used by a dequeuer holding a ticket that an enqueuer has
// already drawn but not yet written: marking the slot at once would make
// that enqueuer go round again.  The bound keeps the operation's step count
// constant-bounded (lock-/wait-freedom are unaffected).
constexpr uint32_t kSlotWaitSpins = 48;

// Polite polling: idle consumers that keep finding a queue empty keep
// re-reading its Head/Tail lines, and on MI355X loads of a line that is also
// the target of device-scope atomics slow those atomics sharply (measured:
// 25% of the waves polling the counter's line halve the FAA rate; polling
// another line costs nothing).  Each wave keeps a count of consecutive EMPTY
// results in LDS and sleeps base * 2^min(count, 6) s_sleep units before its
// next poll; a success resets it.  LDS holds no value at launch, so the count
// is masked to 4 bits.
__device__ __forceinline__ uint8_t* idle_streak() {
  __shared__ uint8_t s_idle[16];
  return &s_idle[((threadIdx.x + threadIdx.y * blockDim.x) / kWave) & 15u];
}
__device__ __forceinline__ void idle_pause(uint32_t base) {
  if (!base) return;
  const uint32_t k = *idle_streak() & 15u;
  if (!k) return;
  const uint32_t n = base << (k < 6u ? k : 6u);
  for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(1);
}
__device__ __forceinline__ void idle_note(uint32_t base, bool empty) {
  if (!base) return;
  uint8_t* p = idle_streak();
  const uint32_t k = *p & 15u;
  *p = (uint8_t)(empty ? (k < 15u ? k + 1u : 15u) : 0u);
}

// Busy wait of roughly `cycles` shader clocks, used by the injected-delay
// adversary and by polite polling loops.
__device__ __forceinline__ void spin_cycles(uint64_t cycles) {
  const uint64_t t0 = __builtin_readcyclecounter();
  while (__builtin_readcyclecounter() - t0 < cycles) __builtin_amdgcn_s_sleep(1);
}

}  // namespace gcq
