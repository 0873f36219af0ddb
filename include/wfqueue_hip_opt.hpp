// wfqueue_hip_opt.hpp — drop-in API of the WFQ (Yang & Mellor-Crummey)
// wait-free queue for the reference harnesses (throughout_bench.cpp:180-220,
// correctness/history.cpp:82-104, concurrent_bfs/bfs_bench.cpp:60-78,
// 530-534, concurrent_rt/rt_final.cpp:137-170), backed by this library's
// implementation (gcq_amd/csrc/gcq/ymc.hpp).
//
// Same names and conventions as the reference's wfqueue_hip_opt.hpp:26-58,
// :527-568, :594-764: WF_SEGMENT_SIZE, WF_PATIENCE,
// WF_PREALLOC_OPS_PER_THREAD, WF_EMPTY (= 0: the value domain excludes 0);
// wf_queue / wf_handle; host helpers wf_queue_host_init,
// wf_queue_host_init_for_bfs and wf_queue_reset_for_bfs; device wf_enqueue
// (false once the cell pool is exhausted) and wf_dequeue (WF_EMPTY when
// empty).  The cell pool is one device allocation published as
// wf_queue::pool.segments, which the harnesses free themselves
// (throughout_bench.cpp:195-204).  Cells are 16 bytes (the reference's are
// 24) and the queue needs no atomicMax per access.
#ifndef GCQ_DROPIN_WFQUEUE_HIP_OPT_HPP
#define GCQ_DROPIN_WFQUEUE_HIP_OPT_HPP
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "gcq/ymc.hpp"

#ifndef WF_SEGMENT_SIZE
#define WF_SEGMENT_SIZE 1024u
#endif
#ifndef WF_PATIENCE
#define WF_PATIENCE 16
#endif
#ifndef WF_PREALLOC_OPS_PER_THREAD
#define WF_PREALLOC_OPS_PER_THREAD 512
#endif
#ifndef WF_SEGMENT_SAFETY
#define WF_SEGMENT_SAFETY 8
#endif
#ifndef WF_EMPTY
#define WF_EMPTY 0ull
#endif

struct wf_handle {
  gcq::ymc::Handle h;
};
static_assert(sizeof(wf_handle) == 64, "wf_handle");

struct wf_pool {
  void* segments;         // the cell array (one allocation)
  uint64_t cells;         // cells in it
  uint32_t capacity;      // in WF_SEGMENT_SIZE units
  uint32_t active_segments;
};
struct wf_queue {
  gcq::ymc::Queue q;      // header; cells and handles are external
  wf_pool pool;
};

__device__ inline bool wf_enqueue(wf_queue* q, wf_handle* h, uint64_t v) {
  gcq::ymc::Handle* hs = (gcq::ymc::Handle*)q->q.g.ext_handles;
  return gcq::ymc::enqueue(&q->q, hs, (uint32_t)((gcq::ymc::Handle*)h - hs), v);
}
__device__ inline uint64_t wf_dequeue(wf_queue* q, wf_handle* h) {
  gcq::ymc::Handle* hs = (gcq::ymc::Handle*)q->q.g.ext_handles;
  const uint64_t v = gcq::ymc::dequeue(&q->q, hs, (uint32_t)((gcq::ymc::Handle*)h - hs));
  if (v != WF_EMPTY) gcq::dropin_acquire();
  return v;
}

namespace gcq_dropin {
__global__ void wfq_init_kernel(wf_queue* q, wf_handle* h, void* cells, uint64_t ncells, uint32_t nthreads) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::ymc::init_slice(&q->q, (gcq::ymc::Handle*)h, ncells, nthreads, WF_PATIENCE, gid, st);
  if (gid == 0) {
    q->q.g.ext_cells = (uint64_t)cells;
    q->q.g.ext_handles = (uint64_t)h;
    q->pool.segments = cells;
    q->pool.cells = ncells;
    q->pool.capacity = (uint32_t)((ncells + WF_SEGMENT_SIZE - 1) / WF_SEGMENT_SIZE);
    q->pool.active_segments = 0;
  }
}
inline void wfq_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
    std::abort();
  }
}
inline void wfq_create(wf_queue** d_q, wf_handle** d_h, int nthreads, uint64_t ncells) {
  if (nthreads <= 0) {
    std::fprintf(stderr, "wf_queue_host_init: nthreads must be > 0\n");
    std::abort();
  }
  void* cells = nullptr;
  wfq_check(hipMalloc((void**)d_q, sizeof(wf_queue)), "hipMalloc(wf_queue)");
  wfq_check(hipMalloc((void**)d_h, sizeof(wf_handle) * (size_t)nthreads), "hipMalloc(wf_handle[])");
  wfq_check(hipMalloc(&cells, 16ull * ncells), "hipMalloc(cell pool)");
  wfq_check(hipMemset(cells, 0, 16ull * ncells), "hipMemset(cell pool)");
  hipLaunchKernelGGL(wfq_init_kernel, dim3((nthreads + 255) / 256), dim3(256), 0, 0, *d_q, *d_h, cells, ncells,
                     (uint32_t)nthreads);
  wfq_check(hipDeviceSynchronize(), "wfq_init_kernel");
}
}  // namespace gcq_dropin

// General pool: the reference's budget of (3 x WF_PREALLOC_OPS_PER_THREAD +
// 128) cells per thread plus WF_SEGMENT_SAFETY segments (:594-618).
inline void wf_queue_host_init(wf_queue** d_q, wf_handle** d_h, int nthreads) {
  const uint64_t ncells = (uint64_t)nthreads * (3ull * WF_PREALLOC_OPS_PER_THREAD + 128ull) +
                          (uint64_t)WF_SEGMENT_SAFETY * WF_SEGMENT_SIZE;
  gcq_dropin::wfq_create(d_q, d_h, nthreads, ncells);
}

// BFS pool (:694-760): the reference's (V + E/chunk) * 1.2 + 10000 cells,
// raised to 2V + T * min(V, 16384) + 2^20.  The BFS harness ping-pongs two
// queues over a whole traversal without rewinding them (bfs_bench.cpp:
// 583-623), and every idle lane can burn one cell per level, so the pool
// covers that burn for up to 16,384 levels (16 B per cell; ~2 GB at 8,192
// threads).
inline void wf_queue_host_init_for_bfs(wf_queue** d_q, wf_handle** d_h, int nthreads, long long E, int V, int chunk) {
  const uint64_t ref = (uint64_t)(((double)V + (double)((E + chunk - 1) / (chunk > 0 ? chunk : 1))) * 1.2 + 10000.0) +
                       (uint64_t)WF_SEGMENT_SAFETY * WF_SEGMENT_SIZE;
  const uint64_t levels = (uint64_t)V < 16384ull ? (uint64_t)V : 16384ull;
  const uint64_t ours = 2ull * (uint64_t)V + (uint64_t)nthreads * levels + (1ull << 20);
  gcq_dropin::wfq_create(d_q, d_h, nthreads, ref > ours ? ref : ours);
}

// Back to empty with the same pool (:646-692).
inline void wf_queue_reset_for_bfs(wf_queue* d_q, wf_handle* d_h, int nthreads) {
  wf_queue hq;
  gcq_dropin::wfq_check(hipMemcpy(&hq, d_q, sizeof hq, hipMemcpyDeviceToHost), "hipMemcpy(wf_queue)");
  gcq_dropin::wfq_check(hipMemset(hq.pool.segments, 0, 16ull * hq.pool.cells), "hipMemset(cell pool)");
  hipLaunchKernelGGL(gcq_dropin::wfq_init_kernel, dim3((nthreads + 255) / 256), dim3(256), 0, 0, d_q, d_h,
                     hq.pool.segments, hq.pool.cells, (uint32_t)nthreads);
  gcq_dropin::wfq_check(hipDeviceSynchronize(), "wfq_init_kernel");
}

#endif  // GCQ_DROPIN_WFQUEUE_HIP_OPT_HPP
