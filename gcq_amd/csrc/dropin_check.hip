// dropin_check.hip — self-check of the drop-in device API headers in
// include/ (gwf_queue.hpp / gwf_ring.hpp, glf_queue.hpp,
// wfqueue_hip_opt.hpp, sfqueue_hip.hpp + .cpp), compiled into the library so
// every build compiles them for gfx950.
//
// For each API it checks the reference conventions a harness relies on:
// objects initialised by one thread as the harnesses do (e.g.
// throughout_bench.cpp:128-131), sequential FIFO order, the exact capacity
// of the bounded queues (Enqueue_Ptr / fq_enqueue_ptr return false when
// full, SFQ's non-blocking enqueue BUSY), the empty sentinels (false,
// wf_entry::BOT, FQ_BOT, WF_EMPTY, SFQ_BUSY), and exactly-once delivery
// with concurrent producers and consumers.
#define FQ_N 4096
#include "../../include/glf_queue.hpp"
#include "../../include/gwf_ring.hpp"
#include "../../include/sfqueue_hip.hpp"
#include "../../include/sfqueue_hip.cpp"
#include "../../include/wfqueue_hip_opt.hpp"
#include "common.hpp"

namespace {

using GCfg = gwf_4k;
constexpr uint32_t kGN = 4096;
using GMpmc = wf_mpmc_t<GCfg, kGN>;
using GRing = wf_t<GCfg, kGN>;
constexpr uint32_t kSpin = 1u << 22;

enum Api { GWF_PTR = 0, GWF_RING = 1, GLF_PTR = 2, GLF_RING = 3, WFQ = 4, SFQ_BLOCK = 5, SFQ_NB = 6 };

// One adapter per API: enq(v) / deq(&v) with v in 1..2^31, the harness
// value conventions (ring APIs carry v - 1 as the index).
struct AGwfPtr {
  GMpmc* q;
  __device__ bool enq(uint32_t v) const { return Enqueue_Ptr<GCfg, kGN>(q, v); }
  __device__ bool deq(uint32_t* v) const {
    uint64_t x = 0;
    if (!Dequeue_Ptr<GCfg, kGN>(q, &x)) return false;
    *v = (uint32_t)x;
    return true;
  }
};
struct AGwfRing {
  GRing* q;
  __device__ bool enq(uint32_t v) const {
    Enqueue_wf<GCfg, kGN>(q, v - 1);
    return true;
  }
  __device__ bool deq(uint32_t* v) const {
    const uint32_t x = Dequeue_wf<GCfg, kGN>(q);
    if (x == wf_entry::BOT) return false;
    *v = x + 1;
    return true;
  }
};
struct AGlfPtr {
  fq_mpmc_t* q;
  __device__ bool enq(uint32_t v) const { return fq_enqueue_ptr(q, v); }
  __device__ bool deq(uint32_t* v) const {
    uint64_t x = 0;
    if (!fq_dequeue_ptr(q, &x)) return false;
    *v = (uint32_t)x;
    return true;
  }
};
struct AGlfRing {
  fq_ring_t* q;
  __device__ bool enq(uint32_t v) const {
    fq_enqueue(q, v - 1);
    return true;
  }
  __device__ bool deq(uint32_t* v) const {
    const uint32_t x = fq_dequeue(q);
    if (x == FQ_BOT) return false;
    *v = x + 1;
    return true;
  }
};
struct AWfq {
  wf_queue* q;
  wf_handle* h;
  __device__ bool enq(uint32_t v) const { return wf_enqueue(q, &h[blockIdx.x * blockDim.x + threadIdx.x], v); }
  __device__ bool deq(uint32_t* v) const {
    const uint64_t x = wf_dequeue(q, &h[blockIdx.x * blockDim.x + threadIdx.x]);
    if (x == WF_EMPTY) return false;
    *v = (uint32_t)x;
    return true;
  }
};
struct ASfqBlock {  // blocking enqueue; the dequeue side uses the non-blocking call (a blocking
  sfq_queue* q;     // dequeue on an empty queue waits for an item by design)
  __device__ bool enq(uint32_t v) const { return sfq_enqueue_blocking_u32(q, v) == SFQ_SUCCESS; }
  __device__ bool deq(uint32_t* v) const { return sfq_dequeue_nb_u32(q, v) == SFQ_SUCCESS; }
};
struct ASfqNb {
  sfq_queue* q;
  __device__ bool enq(uint32_t v) const { return sfq_enqueue_nb_u32(q, v) == SFQ_SUCCESS; }
  __device__ bool deq(uint32_t* v) const { return sfq_dequeue_nb_u32(q, v) == SFQ_SUCCESS; }
};

// errs: 0 order, 1 capacity, 2 not empty at end, 3 duplicate / out of range,
// 4 missing, 5 spin limit
__global__ void k_init(int api, void* q) {
  if (threadIdx.x || blockIdx.x) return;
  switch (api) {
    case GWF_PTR: wf_mpmc_init<GCfg, kGN>((GMpmc*)q); break;
    case GWF_RING: wf_init<GCfg, kGN>((GRing*)q); break;
    case GLF_PTR: fq_init((fq_mpmc_t*)q); break;
    case GLF_RING: fq_ring_init_empty((fq_ring_t*)q); break;
    default: break;
  }
}

// One thread: fill to capacity (or `cap` items), drain in order, see empty.
template <class A>
__global__ void k_sequential(A a, uint32_t cap, bool bounded, uint32_t* errs) {
  if (threadIdx.x || blockIdx.x) return;
  uint32_t n = 0;
  while (n < cap && a.enq(n + 1)) ++n;
  if (n != cap) gcq::or_rlx(&errs[1], 1u);
  if (bounded && a.enq(cap + 1)) gcq::or_rlx(&errs[1], 2u);  // full must refuse
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t v = 0;
    if (!a.deq(&v) || v != i + 1) gcq::faa_rlx(&errs[0], 1u);
  }
  uint32_t v = 0;
  if (a.deq(&v)) gcq::or_rlx(&errs[2], 1u);
}

// Producers (tid < P) enqueue ops items each; everyone else dequeues until
// all P*ops items are out.  Item = producer * ops + seq + 1.
template <class A>
__global__ void k_concurrent(A a, uint32_t producers, uint32_t ops, uint32_t* counts, uint32_t* consumed,
                             uint32_t* errs) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = producers * ops;
  if (tid < producers) {
    for (uint32_t i = 0; i < ops; ++i) {
      uint32_t spins = 0;
      while (!a.enq(tid * ops + i + 1)) {
        if (++spins > kSpin) {
          gcq::or_rlx(&errs[5], 1u);
          return;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    return;
  }
  uint32_t idle = 0;
  while (gcq::ld_rlx(consumed) < total) {
    uint32_t v = 0;
    if (a.deq(&v)) {
      idle = 0;
      if (v == 0 || v > total || gcq::faa_rlx(&counts[v - 1], 1u) != 0u) gcq::or_rlx(&errs[3], 1u);
      gcq::faa_rlx(consumed, 1u);
    } else if (++idle > kSpin) {
      gcq::or_rlx(&errs[5], 2u);
      return;
    } else {
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

__global__ void k_missing(const uint32_t* counts, uint32_t n, uint32_t* errs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && counts[i] != 1u) gcq::or_rlx(&errs[4], 1u);
}

template <class A>
int run_api(A a, uint32_t cap, bool bounded, uint32_t threads, uint32_t producers, uint32_t ops, uint32_t* errs,
            void (*reinit)(void*), void* rarg) {
  uint32_t *counts = nullptr, *consumed = nullptr;
  GCQ_HIP(hipMalloc(&counts, 4ull * producers * ops + 4));
  GCQ_HIP(hipMalloc(&consumed, 4));
  hipLaunchKernelGGL(k_sequential<A>, dim3(1), dim3(64), 0, 0, a, cap, bounded, errs);
  GCQ_HIP(hipDeviceSynchronize());
  reinit(rarg);
  GCQ_HIP(hipMemset(counts, 0, 4ull * producers * ops + 4));
  GCQ_HIP(hipMemset(consumed, 0, 4));
  hipLaunchKernelGGL(k_concurrent<A>, dim3(threads / 64), dim3(64), 0, 0, a, producers, ops, counts, consumed, errs);
  GCQ_HIP(hipDeviceSynchronize());
  const uint32_t n = producers * ops;
  hipLaunchKernelGGL(k_missing, dim3((n + 255) / 256), dim3(256), 0, 0, counts, n, errs);
  GCQ_HIP(hipDeviceSynchronize());
  GCQ_HIP(hipGetLastError());
  (void)hipFree(counts);
  (void)hipFree(consumed);
  return GCQ_OK;
}

struct Reinit {
  int api;
  void* q;
};
void reinit_ring(void* p) {
  Reinit* r = (Reinit*)p;
  hipLaunchKernelGGL(k_init, dim3(1), dim3(1), 0, 0, r->api, r->q);
  (void)hipDeviceSynchronize();
}
struct ReinitWfq {
  wf_queue* q;
  wf_handle* h;
  int threads;
};
void reinit_wfq(void* p) {
  ReinitWfq* r = (ReinitWfq*)p;
  wf_queue_reset_for_bfs(r->q, r->h, r->threads);
}
void reinit_sfq(void* p) {
  sfq_queue* q = (sfq_queue*)p;
  (void)hipMemset(q, 0, sizeof(sfq_queue));  // all-zero is an empty queue (bfs_bench.cpp:538-542)
  (void)hipDeviceSynchronize();
}

}  // namespace

// Runs the drop-in API `api` (0 G-WFQ Enqueue_Ptr/Dequeue_Ptr, 1 G-WFQ
// Enqueue_wf/Dequeue_wf, 2 G-LFQ fq_*_ptr, 3 G-LFQ fq_enqueue/fq_dequeue,
// 4 WFQ, 5 SFQ blocking enqueue, 6 SFQ non-blocking) with `threads`
// threads; errs_out receives the 6 error counters described above.
extern "C" int gcq_dropin_selfcheck(int32_t api, uint32_t threads, uint32_t ops, uint32_t* errs_out) {
  // thread ids must stay below gwf_4k::MAX_THREADS = 4095
  GCQ_REQUIRE(errs_out && threads >= 128 && threads % 64 == 0 && threads <= 4032 && ops >= 1,
              "threads must be a multiple of 64 in [128, 4032]");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  uint32_t* errs = nullptr;
  GCQ_HIP(hipMalloc(&errs, 32));
  GCQ_HIP(hipMemset(errs, 0, 32));
  const uint32_t producers = threads / 4;
  int rc = GCQ_OK;
  void* q = nullptr;
  switch (api) {
    case GWF_PTR:
    case GWF_RING:
    case GLF_PTR:
    case GLF_RING: {
      const size_t bytes = api == GWF_PTR ? sizeof(GMpmc)
                           : api == GWF_RING ? sizeof(GRing)
                           : api == GLF_PTR ? sizeof(fq_mpmc_t)
                                            : sizeof(fq_ring_t);
      GCQ_HIP(hipMalloc(&q, bytes));
      Reinit r{api, q};
      reinit_ring(&r);
      // ring APIs carry indices 0..N-1: every index at most once in flight
      const uint32_t cap = api <= GWF_RING ? kGN : FQ_N;
      const uint32_t ops_c = (api == GWF_RING || api == GLF_RING) ? std::min(ops, cap / producers) : ops;
      if (api == GWF_PTR)
        rc = run_api(AGwfPtr{(GMpmc*)q}, cap, true, threads, producers, ops_c, errs, reinit_ring, &r);
      else if (api == GWF_RING)
        rc = run_api(AGwfRing{(GRing*)q}, cap, false, threads, producers, ops_c, errs, reinit_ring, &r);
      else if (api == GLF_PTR)
        rc = run_api(AGlfPtr{(fq_mpmc_t*)q}, cap, true, threads, producers, ops_c, errs, reinit_ring, &r);
      else
        rc = run_api(AGlfRing{(fq_ring_t*)q}, cap, false, threads, producers, ops_c, errs, reinit_ring, &r);
      (void)hipFree(q);
      break;
    }
    case WFQ: {
      wf_queue* wq = nullptr;
      wf_handle* wh = nullptr;
      wf_queue_host_init(&wq, &wh, (int)threads);
      ReinitWfq r{wq, wh, (int)threads};
      rc = run_api(AWfq{wq, wh}, 1024, false, threads, producers, ops, errs, reinit_wfq, &r);
      wf_queue hq;
      (void)hipMemcpy(&hq, wq, sizeof hq, hipMemcpyDeviceToHost);
      (void)hipFree(hq.pool.segments);
      (void)hipFree(wq);
      (void)hipFree(wh);
      break;
    }
    case SFQ_BLOCK:
    case SFQ_NB: {
      sfq_queue* sq = nullptr;
      sfq_handle* sh = nullptr;
      sfq_queue_host_init(&sq, &sh, (int)threads);
      if (api == SFQ_BLOCK)
        rc = run_api(ASfqBlock{sq}, SFQ_QUEUE_LENGTH, false, threads, producers, ops, errs, reinit_sfq, sq);
      else
        rc = run_api(ASfqNb{sq}, SFQ_QUEUE_LENGTH, true, threads, producers, ops, errs, reinit_sfq, sq);
      sfq_queue_destroy(sq, sh);
      break;
    }
    default:
      (void)hipFree(errs);
      GCQ_REQUIRE(false, "unknown drop-in API %d", api);
  }
  if (rc == GCQ_OK) GCQ_HIP(hipMemcpy(errs_out, errs, 24, hipMemcpyDeviceToHost));
  (void)hipFree(errs);
  return rc;
}
