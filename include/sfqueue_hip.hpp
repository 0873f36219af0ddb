// sfqueue_hip.hpp — drop-in API of the SFQ (scalable FIFO queue) for the
// reference harnesses (throughout_bench.cpp:224-292, correctness/history.cpp:
// 106-140, concurrent_bfs/bfs_bench.cpp:94-132, 535-542,
// concurrent_rt/rt_final.cpp:172-298).
//
// Same names, constants and conventions as the reference's
// sfqueue_hip.hpp:13-104 and sfqueue_hip.cpp:13-180, 356-399:
// SFQ_QUEUE_LENGTH, SFQ_MAX_THREADS, SFQ_BACKOFF_ITERS, SFQ_EMPTY,
// SFQ_SUCCESS / SFQ_BUSY / SFQ_CLOSED; struct sfq_queue with the fields the
// harnesses touch directly (head, tail, nprocs, items[], ids[]; an all-zero
// object is an empty queue), sfq_handle; blocking and non-blocking u32
// operations, the 64-bit compatibility wrappers, init helpers and host
// init / destroy.  Own design: head and tail are 64-bit tickets on their own
// 128-B lines (so the 32-bit lap-id invariant of the reference,
// sfqueue_hip.hpp:41-51, does not cap the thread count), and item / id
// hand-off uses device-scope loads and stores ordered by s_waitcnt.
#ifndef GCQ_DROPIN_SFQUEUE_HIP_HPP
#define GCQ_DROPIN_SFQUEUE_HIP_HPP
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "gcq/prim.hpp"

#ifndef SFQ_QUEUE_LENGTH
#define SFQ_QUEUE_LENGTH 65536u
#endif
#ifndef SFQ_MAX_THREADS
#define SFQ_MAX_THREADS 65535u
#endif
#ifndef SFQ_BACKOFF_ITERS
#define SFQ_BACKOFF_ITERS 2048
#endif
#define SFQ_EMPTY 0ull
#define SFQ_SUCCESS 0
#define SFQ_BUSY 1
#define SFQ_CLOSED 2
static_assert((SFQ_QUEUE_LENGTH & (SFQ_QUEUE_LENGTH - 1u)) == 0, "SFQ_QUEUE_LENGTH must be a power of two");

struct sfq_queue {
  alignas(128) uint64_t head;  // dequeue tickets
  uint64_t pad0[15];
  uint64_t tail;               // enqueue tickets
  uint64_t pad1[15];
  uint32_t closed;
  uint32_t nprocs;
  uint32_t pad2[30];
  uint32_t items[SFQ_QUEUE_LENGTH];
  uint32_t ids[SFQ_QUEUE_LENGTH];  // 2 * lap: free for that lap's enqueue; + 1: full
};

struct sfq_handle {
  uint32_t thread_id;
  uint32_t dummy[15];
};

namespace gcq_dropin {
__device__ __forceinline__ uint32_t sfq_slot(uint64_t t) { return (uint32_t)(t & (SFQ_QUEUE_LENGTH - 1u)); }
__device__ __forceinline__ uint32_t sfq_lap_id(uint64_t t) { return (uint32_t)((t / SFQ_QUEUE_LENGTH) * 2u); }
__device__ __forceinline__ void sfq_backoff() {
  for (int i = 0; i < SFQ_BACKOFF_ITERS / 64; ++i) __builtin_amdgcn_s_sleep(1);
}
__device__ __forceinline__ void sfq_fill(sfq_queue* q, uint32_t j, uint32_t id, uint32_t item) {
  gcq::st_rlx(&q->items[j], item);
  gcq::vm_wait();  // the item before the id that publishes it
  gcq::st_rlx(&q->ids[j], id + 1u);
}
__device__ __forceinline__ uint32_t sfq_take(sfq_queue* q, uint32_t j, uint32_t id) {
  const uint32_t item = gcq::ld_rlx(&q->items[j]);
  gcq::vm_wait();  // the item read before the slot is released to the next lap
  gcq::st_rlx(&q->ids[j], id + 2u);
  return item;
}
// Non-blocking claim on *ctr for the wave's active lanes: lane r tests ticket
// t0 + r (t0 read once by the leader); one CAS claims the ready prefix
// [t0, t0 + k).  Returns the lane's ticket, or ~0 when its op is BUSY.  The
// reference has every lane CAS the same counter value, so at most one lane
// of all contenders wins per round (sfqueue_hip.cpp:89-128).
__device__ __forceinline__ uint64_t sfq_nb_claim(uint64_t* ctr, const uint32_t* ids, uint32_t add) {
  const uint64_t m = gcq::active_mask();
  const uint32_t leader = gcq::first_lane(m);
  uint64_t t0 = 0;
  if (gcq::lane_id() == leader) t0 = gcq::ld_rlx(ctr);
  t0 = gcq::bcast64(t0, leader);
  const uint32_t rank = gcq::rank_in(m);
  const uint64_t t = t0 + rank;
  const bool ready = gcq::ld_rlx(&ids[sfq_slot(t)]) == sfq_lap_id(t) + add;
  const uint64_t not_ready = m & ~__ballot(ready);
  const uint32_t k = (uint32_t)__builtin_popcountll(not_ready ? (m & ((1ull << gcq::first_lane(not_ready)) - 1ull)) : m);
  uint32_t won = 0;
  if (gcq::lane_id() == leader && k > 0) {
    uint64_t expect = t0;
    won = gcq::cas_rlx(ctr, expect, t0 + k) ? 1u : 0u;
  }
  won = gcq::bcast32(won, leader);
  return (won && rank < k) ? t : ~0ull;
}
}  // namespace gcq_dropin

__device__ inline int sfq_is_closed(const sfq_queue* q) { return gcq::ld_rlx(&q->closed) != 0u; }
__device__ inline void sfq_close(sfq_queue* q) { gcq::st_rlx(&q->closed, 1u); }

__device__ inline int sfq_enqueue_blocking_u32(sfq_queue* q, uint32_t item) {
  if (sfq_is_closed(q)) return SFQ_CLOSED;
  const uint64_t t = gcq::wave_faa(&q->tail, (uint64_t)1);  // one FAA per wave
  const uint32_t j = gcq_dropin::sfq_slot(t), id = gcq_dropin::sfq_lap_id(t);
  while (gcq::ld_rlx(&q->ids[j]) != id) gcq_dropin::sfq_backoff();
  gcq_dropin::sfq_fill(q, j, id, item);
  return SFQ_SUCCESS;
}
__device__ inline int sfq_dequeue_blocking_u32(sfq_queue* q, uint32_t* out) {
  if (sfq_is_closed(q)) return SFQ_CLOSED;
  const uint64_t h = gcq::wave_faa(&q->head, (uint64_t)1);  // one FAA per wave
  const uint32_t j = gcq_dropin::sfq_slot(h), id = gcq_dropin::sfq_lap_id(h);
  while (gcq::ld_rlx(&q->ids[j]) != id + 1u) gcq_dropin::sfq_backoff();
  *out = gcq_dropin::sfq_take(q, j, id);
  gcq::dropin_acquire();
  return SFQ_SUCCESS;
}
// Non-blocking: take the ticket only if its slot is ready (else BUSY).
__device__ inline int sfq_enqueue_nb_u32(sfq_queue* q, uint32_t item) {
  if (sfq_is_closed(q)) return SFQ_CLOSED;
  const uint64_t t = gcq_dropin::sfq_nb_claim(&q->tail, q->ids, 0u);
  if (t == ~0ull) return SFQ_BUSY;
  gcq_dropin::sfq_fill(q, gcq_dropin::sfq_slot(t), gcq_dropin::sfq_lap_id(t), item);
  return SFQ_SUCCESS;
}
__device__ inline int sfq_dequeue_nb_u32(sfq_queue* q, uint32_t* out) {
  if (sfq_is_closed(q)) return SFQ_CLOSED;
  const uint64_t h = gcq_dropin::sfq_nb_claim(&q->head, q->ids, 1u);
  if (h == ~0ull) return SFQ_BUSY;
  *out = gcq_dropin::sfq_take(q, gcq_dropin::sfq_slot(h), gcq_dropin::sfq_lap_id(h));
  gcq::dropin_acquire();
  return SFQ_SUCCESS;
}

// 64-bit compatibility wrappers: item 0 becomes 1 so SFQ_EMPTY stays free.
__device__ inline void sfq_enqueue(sfq_queue* q, sfq_handle*, uint64_t v) {
  uint32_t item = (uint32_t)(v & 0xFFFFFFFFu);
  if (item == 0u) item = 1u;
  (void)sfq_enqueue_blocking_u32(q, item);
}
__device__ inline uint64_t sfq_dequeue(sfq_queue* q, sfq_handle*) {
  uint32_t item = 0u;
  return sfq_dequeue_blocking_u32(q, &item) == SFQ_SUCCESS ? (uint64_t)item : SFQ_EMPTY;
}

__device__ inline void sfq_queue_init(sfq_queue* q, uint32_t nprocs) {
  q->head = 0;
  q->tail = 0;
  q->closed = 0;
  q->nprocs = nprocs;
}
__device__ inline void sfq_handle_init(sfq_handle* h, uint32_t tid) {
  h->thread_id = tid;
  for (int i = 0; i < 15; ++i) h->dummy[i] = 0u;
}

static __global__ void sfq_init_kernel(sfq_queue* q, sfq_handle* handles, int num_threads) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, total = blockDim.x * gridDim.x;
  if (tid == 0) sfq_queue_init(q, (uint32_t)num_threads);
  for (int i = tid; i < (int)SFQ_QUEUE_LENGTH; i += total) {
    q->items[i] = 0u;
    q->ids[i] = 0u;
  }
  for (int i = tid; i < num_threads; i += total) sfq_handle_init(&handles[i], (uint32_t)i);
}

inline void sfq_queue_host_init(sfq_queue** d_q, sfq_handle** d_handles, int num_threads) {
  if (num_threads < 0) {
    std::fprintf(stderr, "sfq_queue_host_init: num_threads must be >= 0\n");
    std::abort();
  }
  hipError_t e = hipMalloc((void**)d_q, sizeof(sfq_queue));
  if (e == hipSuccess) e = hipMalloc((void**)d_handles, (size_t)(num_threads > 0 ? num_threads : 1) * sizeof(sfq_handle));
  if (e != hipSuccess) {
    std::fprintf(stderr, "sfq_queue_host_init: %s\n", hipGetErrorString(e));
    std::abort();
  }
  const int blocks = (int)((SFQ_QUEUE_LENGTH + 255u) / 256u);
  hipLaunchKernelGGL(sfq_init_kernel, dim3(blocks), dim3(256), 0, 0, *d_q, *d_handles, num_threads);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    std::fprintf(stderr, "sfq_init_kernel failed: %s\n", hipGetErrorString(e));
    std::abort();
  }
}
inline void sfq_queue_destroy(sfq_queue* d_q, sfq_handle* d_h) {
  if (d_q) (void)hipFree(d_q);
  if (d_h) (void)hipFree(d_h);
}

#endif  // GCQ_DROPIN_SFQUEUE_HIP_HPP
