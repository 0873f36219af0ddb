// glf_queue.hpp — drop-in device API of the G-LFQ lock-free queue for the
// reference harnesses (throughout_bench.cpp:148-178, correctness/history.cpp:
// 65-80, concurrent_bfs/bfs_bench.cpp:48-58, concurrent_rt/rt_final.cpp:
// 53-86), backed by this library's gfx950 ring (gcq_amd/csrc/gcq/glfq.hpp).
//
// Same names and conventions as the reference's glf_queue.hpp:30-51,
// :126-138, :418-440, :467-582: FQ_N (defined by the includer before this
// header, default 32768), FQ_SLOTS, FQ_CYCLE_BITS, FQ_BOT; fq_ring_t and
// fq_mpmc_t (objects the host allocates with sizeof and a single-thread
// kernel initialises), fq_ring_init_empty / fq_init, fq_enqueue /
// fq_dequeue (index, FQ_BOT when empty), fq_enqueue_ptr / fq_dequeue_ptr
// (false = full / empty).  Layout is this library's: Tail, Head and
// Threshold on separate 128-B lines.
#ifndef GCQ_DROPIN_GLF_QUEUE_HPP
#define GCQ_DROPIN_GLF_QUEUE_HPP
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gcq/glfq.hpp"

#ifndef FQ_N
#define FQ_N 32768
#endif
#ifndef FQ_CYCLE_BITS
#define FQ_CYCLE_BITS 8
#endif
static_assert((FQ_N & (FQ_N - 1)) == 0 && FQ_N >= 64, "FQ_N must be a power of two >= 64");
static_assert(FQ_CYCLE_BITS == 8, "the ring keeps an 8-bit cycle tag");
static constexpr uint32_t FQ_SLOTS = 2 * FQ_N;
static constexpr int32_t FQ_THRESH_MAX = 3 * (int32_t)FQ_N - 1;
static constexpr uint32_t FQ_BOT = 0xFFFFFFFEu;  // fq_dequeue's "empty" (glf_queue.hpp:48)

struct alignas(256) fq_ring_t {
  unsigned char bytes[gcq::lfq::ring_bytes(FQ_N)];
};
struct fq_mpmc_t {
  fq_ring_t aq;         // published indices
  fq_ring_t fq;         // free indices
  uint64_t data[FQ_N];  // payloads
};

__device__ inline void fq_ring_init_empty(fq_ring_t* r) {
  gcq::lfq::init_slice((gcq::lfq::Ring*)r, FQ_N, false, 0u, 0u, 1u, 0u);
}
__device__ inline void fq_ring_init_full(fq_ring_t* r) {
  gcq::lfq::init_slice((gcq::lfq::Ring*)r, FQ_N, true, 0u, 0u, 1u, 0u);
}
// aq empty, fq holding every index 0..FQ_N-1; one thread (glf_queue.hpp:437-440).
__device__ inline void fq_init(fq_mpmc_t* q) {
  fq_ring_init_empty(&q->aq);
  fq_ring_init_full(&q->fq);
}

__device__ inline void fq_enqueue(fq_ring_t* r, uint32_t index) { gcq::lfq::enqueue((gcq::lfq::Ring*)r, index); }
__device__ inline uint32_t fq_dequeue(fq_ring_t* r) {
  const uint32_t x = gcq::lfq::dequeue((gcq::lfq::Ring*)r, (gcq::lfq::NoProbe*)nullptr, true);
  if (x == gcq::lfq::kBot) return FQ_BOT;
  gcq::dropin_acquire();
  return x;
}

__device__ inline bool fq_enqueue_ptr(fq_mpmc_t* q, uint64_t v) {
  return gcq::lfq::enqueue_ptr(gcq::lfq::Mpmc{(gcq::lfq::Ring*)&q->aq, (gcq::lfq::Ring*)&q->fq, q->data}, v);
}
__device__ inline bool fq_dequeue_ptr(fq_mpmc_t* q, uint64_t* out) {
  // The reference fences after taking the value (glf_queue.hpp:579), which
  // also orders the caller's later loads after the dequeue.
  if (!gcq::lfq::dequeue_ptr(gcq::lfq::Mpmc{(gcq::lfq::Ring*)&q->aq, (gcq::lfq::Ring*)&q->fq, q->data}, out))
    return false;
  gcq::dropin_acquire();
  return true;
}

#endif  // GCQ_DROPIN_GLF_QUEUE_HPP
