// bfs.hip — queue-driven level-synchronous BFS on an int32 CSR in HBM.
//
// Semantics follow the reference harness (bfs_bench.cpp:327-633): two
// frontier queues ping-pong between levels; every dequeued vertex u has its
// neighbours w tested with CAS(level[w], -1, L+1) and the winner enqueues w;
// the traversal ends when a level discovers nothing.  Depths are therefore
// exact BFS depths (bit-exact against a serial BFS).
//
// MI355X design (own):
//  * frontier queues hold vertex ids directly in the ring (G-WFQ / G-LFQ
//    ring-level API, SFQ items, WFQ values + 1): one ring operation per
//    vertex instead of the reference's two-ring indirection plus fence;
//    ring capacity >= V so no vertex is dropped (the reference's 32,768-slot
//    rings drop frontier vertices silently once a level exceeds them,
//    bfs_bench.cpp:38-39, 380);
//  * expansion is wave-cooperative for every degree: the wave prefix-sums the
//    degrees of the (up to 64) vertices it dequeued and strides its 64 lanes
//    over their concatenated adjacency (load balanced; no separate hub
//    kernel, bfs_bench.cpp:434-496);
//  * visited vertices are filtered with a plain load before the CAS;
//  * newly discovered vertices are staged per wave in LDS and enqueued 64 at
//    a time (one aggregated ticket FAA per 64 vertices);
//  * GCQ_BFS_PERSISTENT runs all levels in one launch with a device-side grid
//    barrier (no host round trip per level; the 4900^2 lattice has ~10^4
//    levels).  Every block is co-resident (checked on the host) and every
//    barrier wait has a spin limit that aborts the run instead of hanging.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "graph.hpp"
#include "queue_impl.hpp"

// ======================================================================== graph handle
struct BfsCtx {
  int kind = -1;
  uint32_t threads = 0, capacity = 0;
  uint64_t ymc_cells = 0;
  int nq = 0;                          // frontier queues (2, or 3 for WFQ)
  void* q[3] = {nullptr, nullptr, nullptr};
  void* h[3] = {nullptr, nullptr, nullptr};  // WFQ handles
  int32_t* level = nullptr;
  void* ctl = nullptr;
  uint32_t* pinned = nullptr;  // host-visible counters for the host-loop mode
};

struct gcq_graph {
  gcqg::Csr host;
  int32_t* d_off = nullptr;
  int32_t* d_idx = nullptr;
  hipStream_t stream = nullptr;
  BfsCtx ctx;
};

namespace bfsk {

using gcq::lane_id;

// Control block: each contended word on its own 128-B line.
struct Ctl {
  uint32_t bar_count, p0[31];
  uint32_t bar_gen, p1[31];
  uint32_t next[3], p2[29];
  uint32_t abort, err, p3[30];
  uint64_t edges, reached, p4[14];
  int32_t levels, p5[31];
};
constexpr uint64_t kBarrierSpinLimit = 1ull << 25;

// ---------------------------------------------------------------- frontier adapters
struct FGwf {
  static constexpr int kQueues = 2;
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ void rewind() const {}
  gcq::gwf::Ring* r;
  __device__ __forceinline__ bool push(uint32_t tid, uint32_t v) const {
    gcq::gwf::NoProbe np;
    gcq::gwf::RingOps<> o(r, tid, &np);
    o.enqueue(v);
    return true;
  }
  __device__ __forceinline__ bool pop(uint32_t tid, uint32_t* v) const {
    gcq::gwf::NoProbe np;
    gcq::gwf::RingOps<> o(r, tid, &np);
    const uint32_t x = o.dequeue(true);
    if (x == gcq::gwf::kBot) return false;
    *v = x;
    return true;
  }
};
struct FLfq {
  static constexpr int kQueues = 2;
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ void rewind() const {}
  gcq::lfq::Ring* r;
  __device__ __forceinline__ bool push(uint32_t, uint32_t v) const {
    gcq::lfq::enqueue(r, v);
    return true;
  }
  __device__ __forceinline__ bool pop(uint32_t, uint32_t* v) const {
    const uint32_t x = gcq::lfq::dequeue(r, (gcq::lfq::NoProbe*)nullptr, true);
    if (x == gcq::lfq::kBot) return false;
    *v = x;
    return true;
  }
};
struct FSfq {
  static constexpr int kQueues = 2;
  __device__ __forceinline__ void recycle(uint64_t, uint64_t, uint32_t) const {}
  __device__ __forceinline__ void rewind() const {}
  gcq::sfq::Queue* q;
  __device__ __forceinline__ bool push(uint32_t, uint32_t v) const {
    return gcq::sfq::enqueue_blocking(q, (uint64_t)v, 1ull << 22) == gcq::sfq::kSuccess;
  }
  __device__ __forceinline__ bool pop(uint32_t, uint32_t* v) const {
    uint64_t x = 0;
    if (!gcq::sfq::dequeue_or_empty(q, &x, 1ull << 20)) return false;
    *v = (uint32_t)x;
    return true;
  }
};
// WFQ cells are never reused by the algorithm, and every level burns up to one
// cell per idle lane, so BFS rotates three WFQ frontiers: while one is the
// level's input and one its output, the third (drained a level earlier)
// is cleared by the grid and rewound at the level's barrier.
struct FYmc {  // values are v + 1: 0 is the queue's empty cell (wfqueue_hip_opt.hpp:42-46)
  static constexpr int kQueues = 3;
  gcq::ymc::Queue* q;
  gcq::ymc::Handle* h;
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

__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct LevelAcc {
  uint64_t edges;
  uint32_t disc;
};

// Enqueue the first n staged vertices (n <= 64) and shift the rest down.
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

// Drain qin (level L) and enqueue the vertices discovered at depth L+1.
// All 64 lanes of the wave must call it together (threads % 64 == 0).
template <class F>
__device__ void expand_level(const F& qin, const F& qout, const int32_t* __restrict__ offs,
                             const int32_t* __restrict__ idx, int32_t* level, int32_t L, uint32_t tid, uint32_t* stage,
                             LevelAcc& acc, uint32_t* err) {
  const uint32_t lane = lane_id();
  uint32_t sc = 0;  // staged vertices (wave-uniform)
  for (;;) {
    uint32_t u = 0;
    const bool got = qin.pop(tid, &u);
    if (__ballot(got) == 0) break;  // the whole wave saw the level's queue empty
    int32_t beg = 0, deg = 0;
    if (got) {
      beg = offs[u];
      deg = offs[u + 1] - beg;
    }
    acc.edges += (uint64_t)deg;
    int32_t incl = deg;  // inclusive prefix of degrees across the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += t;
    }
    const int32_t excl = incl - deg;
    const int32_t total = __shfl(incl, 63, 64);
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t e = base + (int32_t)lane;
      const int32_t ec = e < total ? e : total - 1;
      int lo = 0, hi = 63;  // owner = first lane whose inclusive prefix exceeds ec
      for (int s = 0; s < 6; ++s) {
        const int mid = (lo + hi) >> 1;
        const int32_t pv = __shfl(incl, mid, 64);
        if (pv > ec) hi = mid;
        else lo = mid + 1;
      }
      const int32_t obeg = __shfl(beg, lo, 64);
      const int32_t oexcl = __shfl(excl, lo, 64);
      bool disc = false;
      uint32_t w = 0;
      if (e < total) {
        w = (uint32_t)idx[obeg + (ec - oexcl)];
        if (gcq::ld_rlx(&level[w]) == -1) {
          int32_t expect = -1;
          disc = gcq::cas_rlx(&level[w], expect, L + 1);
        }
      }
      const uint64_t dm = __ballot(disc);
      if (disc) stage[sc + gcq::rank_in(dm)] = w;
      sc += (uint32_t)__builtin_popcountll(dm);
      acc.disc += disc ? 1u : 0u;
      if (sc >= 64) flush(qout, tid, stage, sc, 64, err);
    }
  }
  if (sc) flush(qout, tid, stage, sc, sc, err);
}

// Sense-reversing grid barrier over co-resident blocks.  The last block to
// arrive runs `last()` before releasing the others.  Returns false once any
// block has aborted (spin limit), so every block leaves its loop.
template <class Last>
__device__ bool grid_barrier(Ctl* c, uint32_t nblocks, Last last) {
  gcq::vm_wait();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = gcq::ld_rlx(&c->bar_gen);
    gcq::vm_wait();  // generation read before arriving
    const uint32_t a = gcq::faa_rlx(&c->bar_count, 1u);
    if (a == nblocks - 1) {
      last();
      gcq::st_rlx(&c->bar_count, 0u);
      gcq::vm_wait();
      gcq::st_rlx(&c->bar_gen, g + 1u);
    } else {
      uint64_t spins = 0;
      while (gcq::ld_rlx(&c->bar_gen) == g) {
        if (gcq::ld_rlx(&c->abort)) break;
        if (++spins > kBarrierSpinLimit) {
          gcq::st_rlx(&c->abort, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
  }
  __syncthreads();
  return gcq::ld_rlx(&c->abort) == 0u;
}

template <class F>
__device__ void finish_counts(LevelAcc& acc, uint64_t disc_total, Ctl* c) {
  const uint64_t e = gcq::wave_sum(acc.edges);
  const uint64_t d = gcq::wave_sum(disc_total);
  if (lane_id() == 0) {
    if (e) gcq::faa_rlx(&c->edges, e);
    if (d) gcq::faa_rlx(&c->reached, d);
  }
}

// Frontier queues of one traversal: level L reads q[L % K] and writes
// q[(L + 1) % K]; with K = 3 the third one is recycled during the level.
template <class F>
struct Frontiers {
  F q[3];
  __device__ __forceinline__ const F& in(int32_t L) const { return q[L % F::kQueues]; }
  __device__ __forceinline__ const F& out(int32_t L) const { return q[(L + 1) % F::kQueues]; }
  __device__ __forceinline__ const F& idle(int32_t L) const { return q[(L + 2) % 3]; }
};

template <class F>
__global__ void k_persistent(Frontiers<F> fq, const int32_t* __restrict__ offs, const int32_t* __restrict__ idx,
                             int32_t* level, Ctl* c) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  uint32_t* st = stage[(threadIdx.x / 64) & 15];
  LevelAcc acc{0, 0};
  uint64_t disc_total = 0;
  int32_t L = 0;
  for (;; ++L) {
    if (F::kQueues == 3) fq.idle(L).recycle(tid, nthreads, tid);
    acc.disc = 0;
    expand_level(fq.in(L), fq.out(L), offs, idx, level, L, tid, st, acc, &c->err);
    disc_total += acc.disc;
    const uint32_t d = wave_sum32(acc.disc);
    if (lane_id() == 0 && d) gcq::faa_rlx(&c->next[L % 3], d);
    if (!grid_barrier(c, gridDim.x, [&]() {
          if (F::kQueues == 3) fq.idle(L).rewind();
        }))
      break;
    const uint32_t n = gcq::ld_rlx(&c->next[L % 3]);
    if (blockIdx.x == 0 && threadIdx.x == 0) gcq::st_rlx(&c->next[(L + 2) % 3], 0u);
    if (n == 0) break;
  }
  finish_counts<F>(acc, disc_total, c);
  if (tid == 0) c->levels = L + 1;
}

template <class F>
__global__ void k_level(Frontiers<F> fq, const int32_t* __restrict__ offs, const int32_t* __restrict__ idx,
                        int32_t* level, int32_t L, Ctl* c) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) gcq::st_rlx(&c->next[(L + 1) & 1], 0u);  // the next level's counter
  if (F::kQueues == 3) fq.idle(L).recycle(tid, (uint64_t)gridDim.x * blockDim.x, tid);
  LevelAcc acc{0, 0};
  expand_level(fq.in(L), fq.out(L), offs, idx, level, L, tid, stage[(threadIdx.x / 64) & 15], acc, &c->err);
  const uint32_t d = wave_sum32(acc.disc);
  if (lane_id() == 0 && d) gcq::faa_rlx(&c->next[L & 1], d);
  finish_counts<F>(acc, acc.disc, c);
}

// Host-loop mode: rewind the recycled frontier after its level's launch.
template <class F>
__global__ void k_rewind(Frontiers<F> fq, int32_t L) {
  if (threadIdx.x == 0) fq.idle(L).rewind();
}

// ---------------------------------------------------------------- resets
__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < n; i += st) p[i] = v;
}
__global__ void k_zero_u64(uint64_t* p, uint64_t n) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < n; i += st) p[i] = 0;
}
__global__ void k_reset_gwf(void* a, void* b, gcq::gwf::InitParams p) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::gwf::init_slice((gcq::gwf::Ring*)a, p, false, gid, st);
  gcq::gwf::init_slice((gcq::gwf::Ring*)b, p, false, gid, st);
}
__global__ void k_reset_lfq(void* a, void* b, uint32_t n) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::lfq::init_slice((gcq::lfq::Ring*)a, n, false, 0, gid, st, 0);
  gcq::lfq::init_slice((gcq::lfq::Ring*)b, n, false, 0, gid, st, 0);
}
__global__ void k_reset_sfq(void* a, void* b, uint64_t n) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::sfq::init_slice((gcq::sfq::Queue*)a, n, 0, gid, st);
  gcq::sfq::init_slice((gcq::sfq::Queue*)b, n, 0, gid, st);
}
__global__ void k_reset_ymc(void* q, void* h, uint64_t cells, uint32_t nthreads) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::ymc::init_slice((gcq::ymc::Queue*)q, (gcq::ymc::Handle*)h, cells, nthreads, 16, gid, st);
}
// Seed the level-0 frontier with the source.
template <class F>
__global__ void k_seed(F q, int32_t* level, int32_t src, Ctl* c) {
  if (threadIdx.x == 0) {
    level[src] = 0;
    if (!q.push(0, (uint32_t)src)) c->err = 1;
  }
}

}  // namespace bfsk

// ======================================================================== host side
namespace {

uint32_t pow2_at_least(uint64_t x) {
  uint64_t p = 64;
  while (p < x) p <<= 1;
  return (uint32_t)std::min<uint64_t>(p, 1ull << 31);
}

int ctx_free(gcq_graph* g) {
  BfsCtx& c = g->ctx;
  for (void* p : {c.q[0], c.q[1], c.q[2], c.h[0], c.h[1], c.h[2], (void*)c.level, c.ctl})
    if (p) (void)hipFree(p);
  if (c.pinned) (void)hipHostFree(c.pinned);
  c = BfsCtx();
  return GCQ_OK;
}

int ctx_ensure(gcq_graph* g, int kind, uint32_t threads) {
  BfsCtx& c = g->ctx;
  if (c.kind == kind && c.threads >= threads && c.level) return GCQ_OK;
  ctx_free(g);
  const uint32_t nv = (uint32_t)g->host.nv;
  uint64_t qbytes = 0;
  c.capacity = pow2_at_least(std::max<uint64_t>(nv, threads));
  switch (kind) {
    case GCQ_GWFQ:
      GCQ_REQUIRE(c.capacity <= (1u << 25), "G-WFQ frontier capacity %u exceeds 2^25 (V too large)", c.capacity);
      qbytes = gcq::gwf::ring_bytes(c.capacity, threads);
      break;
    case GCQ_GLFQ: qbytes = gcq::lfq::ring_bytes(c.capacity); break;
    case GCQ_SFQ: qbytes = gcq::sfq::queue_bytes(c.capacity); break;
    case GCQ_WFQ:
      // One level's cells: its frontier (<= V) plus at most one burned cell
      // per dequeuing lane; the queue is recycled two levels later.
      c.ymc_cells = (uint64_t)nv + 4ull * threads + (1ull << 20);
      qbytes = gcq::ymc::queue_bytes(c.ymc_cells);
      break;
    default: GCQ_REQUIRE(false, "unknown queue kind %d", kind);
  }
  c.nq = kind == GCQ_WFQ ? 3 : 2;
  for (int i = 0; i < c.nq; ++i) {
    GCQ_HIP(hipMalloc(&c.q[i], qbytes));
    if (kind == GCQ_WFQ) GCQ_HIP(hipMalloc(&c.h[i], sizeof(gcq::ymc::Handle) * (size_t)threads));
  }
  GCQ_HIP(hipMalloc(&c.level, 4ull * nv));
  GCQ_HIP(hipMalloc(&c.ctl, sizeof(bfsk::Ctl)));
  GCQ_HIP(hipHostMalloc((void**)&c.pinned, 64));
  c.kind = kind;
  c.threads = threads;
  return GCQ_OK;
}

uint32_t grid_of(uint64_t work) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, 8192)); }

int reset_run(gcq_graph* g, int32_t src) {
  BfsCtx& c = g->ctx;
  const uint32_t nv = (uint32_t)g->host.nv;
  hipLaunchKernelGGL(bfsk::k_fill_i32, dim3(grid_of(nv)), dim3(256), 0, g->stream, c.level, (uint64_t)nv, -1);
  hipLaunchKernelGGL(bfsk::k_zero_u64, dim3(1), dim3(256), 0, g->stream, (uint64_t*)c.ctl,
                     (uint64_t)(sizeof(bfsk::Ctl) / 8));
  switch (c.kind) {
    case GCQ_GWFQ: {
      gcq::gwf::InitParams p{c.capacity, c.threads, 0, 64, 128, 256, 0};
      hipLaunchKernelGGL(bfsk::k_reset_gwf, dim3(grid_of(2ull * c.capacity)), dim3(256), 0, g->stream, c.q[0], c.q[1],
                         p);
      hipLaunchKernelGGL(bfsk::k_seed<bfsk::FGwf>, dim3(1), dim3(64), 0, g->stream,
                         bfsk::FGwf{(gcq::gwf::Ring*)c.q[0]}, c.level, src, (bfsk::Ctl*)c.ctl);
      break;
    }
    case GCQ_GLFQ:
      hipLaunchKernelGGL(bfsk::k_reset_lfq, dim3(grid_of(2ull * c.capacity)), dim3(256), 0, g->stream, c.q[0], c.q[1],
                         c.capacity);
      hipLaunchKernelGGL(bfsk::k_seed<bfsk::FLfq>, dim3(1), dim3(64), 0, g->stream,
                         bfsk::FLfq{(gcq::lfq::Ring*)c.q[0]}, c.level, src, (bfsk::Ctl*)c.ctl);
      break;
    case GCQ_SFQ:
      hipLaunchKernelGGL(bfsk::k_reset_sfq, dim3(grid_of(c.capacity)), dim3(256), 0, g->stream, c.q[0], c.q[1],
                         (uint64_t)c.capacity);
      hipLaunchKernelGGL(bfsk::k_seed<bfsk::FSfq>, dim3(1), dim3(64), 0, g->stream,
                         bfsk::FSfq{(gcq::sfq::Queue*)c.q[0]}, c.level, src, (bfsk::Ctl*)c.ctl);
      break;
    case GCQ_WFQ: {
      const uint64_t words = c.ymc_cells * 2;
      for (int i = 0; i < 3; ++i) {
        hipLaunchKernelGGL(bfsk::k_zero_u64, dim3(grid_of(words)), dim3(256), 0, g->stream,
                           (uint64_t*)((char*)c.q[i] + gcq::ymc::kHdrBytes), words);
        hipLaunchKernelGGL(bfsk::k_reset_ymc, dim3(grid_of(c.threads)), dim3(256), 0, g->stream, c.q[i], c.h[i],
                           c.ymc_cells, c.threads);
      }
      hipLaunchKernelGGL(bfsk::k_seed<bfsk::FYmc>, dim3(1), dim3(64), 0, g->stream,
                         bfsk::FYmc{(gcq::ymc::Queue*)c.q[0], (gcq::ymc::Handle*)c.h[0]}, c.level, src,
                         (bfsk::Ctl*)c.ctl);
      break;
    }
  }
  GCQ_HIP(hipGetLastError());
  return GCQ_OK;
}

template <class F>
int traverse(gcq_graph* g, const bfsk::Frontiers<F>& fq, const gcq_bfs_config* cfg, double* kernel_ms) {
  BfsCtx& c = g->ctx;
  bfsk::Ctl* ctl = (bfsk::Ctl*)c.ctl;
  const uint32_t grid = cfg->threads / cfg->block;
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  int rc = GCQ_OK;
  if (cfg->mode == GCQ_BFS_PERSISTENT) {
    int per_cu = 0, dev = 0, cus = 0;
    GCQ_HIP(hipGetDevice(&dev));
    GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bfsk::k_persistent<F>, cfg->block, 0));
    if ((int64_t)grid > (int64_t)per_cu * cus) {
      gcqh::set_error("persistent BFS needs all %u blocks resident; the device holds %d", grid, per_cu * cus);
      rc = GCQ_EINVAL;
    } else {
      GCQ_HIP(hipEventRecord(e0, g->stream));
      hipLaunchKernelGGL(bfsk::k_persistent<F>, dim3(grid), dim3(cfg->block), 0, g->stream, fq, g->d_off, g->d_idx,
                         c.level, ctl);
      GCQ_HIP(hipEventRecord(e1, g->stream));
    }
  } else {
    // Host level loop (bfs_bench.cpp:583-623): one launch per level, the
    // next frontier size read back after each.
    GCQ_HIP(hipEventRecord(e0, g->stream));
    int32_t L = 0;
    for (;; ++L) {
      hipLaunchKernelGGL(bfsk::k_level<F>, dim3(grid), dim3(cfg->block), 0, g->stream, fq, g->d_off, g->d_idx, c.level,
                         L, ctl);
      if (F::kQueues == 3) hipLaunchKernelGGL(bfsk::k_rewind<F>, dim3(1), dim3(64), 0, g->stream, fq, L);
      GCQ_HIP(hipMemcpyAsync(c.pinned, &ctl->next[L & 1], 4, hipMemcpyDeviceToHost, g->stream));
      GCQ_HIP(hipStreamSynchronize(g->stream));
      if (*(volatile uint32_t*)c.pinned == 0) break;
    }
    GCQ_HIP(hipEventRecord(e1, g->stream));
    c.pinned[1] = (uint32_t)(L + 1);  // levels, as the reference counts them (bfs_bench.cpp:622-631)
    GCQ_HIP(hipMemcpyAsync(&ctl->levels, &c.pinned[1], 4, hipMemcpyHostToDevice, g->stream));
  }
  if (rc == GCQ_OK) {
    GCQ_HIP(hipEventSynchronize(e1));
    GCQ_HIP(hipGetLastError());
    *kernel_ms = gcqh::ms_between(e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Copy the CSR to HBM on first use (graphs can be built and inspected on a
// host without a GPU).
int ensure_uploaded(gcq_graph* g) {
  if (g->d_off) return GCQ_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  const int64_t ne = g->host.ne();
  GCQ_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  GCQ_HIP(hipMalloc(&g->d_off, 4ull * (g->host.nv + 1)));
  GCQ_HIP(hipMalloc(&g->d_idx, 4ull * std::max<int64_t>(ne, 1)));
  GCQ_HIP(hipMemcpy(g->d_off, g->host.offsets.data(), 4ull * (g->host.nv + 1), hipMemcpyHostToDevice));
  if (ne) GCQ_HIP(hipMemcpy(g->d_idx, g->host.indices.data(), 4ull * ne, hipMemcpyHostToDevice));
  return GCQ_OK;
}

int finish_create(gcq_graph* g, gcq_graph** out) {
  *out = g;
  return GCQ_OK;
}

}  // namespace

extern "C" int gcq_graph_lattice(int32_t rows, int32_t cols, double keep, double diag, uint64_t seed,
                                 gcq_graph** out) {
  GCQ_REQUIRE(out, "null out");
  GCQ_REQUIRE(keep >= 0 && keep <= 1 && diag >= 0 && diag <= 1, "probabilities must be in [0,1]");
  gcq_graph* g = new gcq_graph();
  if (!gcqg::lattice(rows, cols, keep, diag, seed, &g->host)) {
    delete g;
    GCQ_REQUIRE(false, "bad lattice size %dx%d", rows, cols);
  }
  return finish_create(g, out);
}

extern "C" int gcq_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                              gcq_graph** out) {
  GCQ_REQUIRE(out, "null out");
  GCQ_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1, "bad R-MAT probabilities");
  gcq_graph* g = new gcq_graph();
  if (!gcqg::rmat(scale, edge_factor, a, b, c, seed, &g->host)) {
    delete g;
    GCQ_REQUIRE(false, "bad R-MAT scale %d / edge factor %d", scale, edge_factor);
  }
  return finish_create(g, out);
}

extern "C" int gcq_graph_from_csr(int32_t nv, int64_t ne, const int32_t* offsets, const int32_t* indices,
                                  gcq_graph** out) {
  GCQ_REQUIRE(out && offsets && (indices || ne == 0) && nv > 0 && ne >= 0 && ne < (1ll << 31), "bad CSR");
  GCQ_REQUIRE(offsets[0] == 0 && offsets[nv] == ne, "offsets must start at 0 and end at ne");
  for (int32_t v = 0; v < nv; ++v)
    GCQ_REQUIRE(offsets[v] <= offsets[v + 1], "offsets decrease at vertex %d", v);
  gcq_graph* g = new gcq_graph();
  g->host.nv = nv;
  g->host.offsets.assign(offsets, offsets + nv + 1);
  g->host.indices.assign(indices, indices + ne);
  for (int64_t i = 0; i < ne; ++i)
    if (indices[i] < 0 || indices[i] >= nv) {
      delete g;
      GCQ_REQUIRE(false, "index %lld out of range", (long long)i);
    }
  return finish_create(g, out);
}

extern "C" int gcq_graph_info(const gcq_graph* g, int32_t* nv, int64_t* ne) {
  GCQ_REQUIRE(g, "null graph");
  if (nv) *nv = g->host.nv;
  if (ne) *ne = g->host.ne();
  return GCQ_OK;
}

extern "C" int gcq_graph_copy_csr(const gcq_graph* g, int32_t* offsets, int32_t* indices) {
  GCQ_REQUIRE(g, "null graph");
  if (offsets) std::memcpy(offsets, g->host.offsets.data(), 4ull * (g->host.nv + 1));
  if (indices) std::memcpy(indices, g->host.indices.data(), 4ull * g->host.ne());
  return GCQ_OK;
}

extern "C" int32_t gcq_graph_pick_source(const gcq_graph* g, uint64_t seed, uint64_t k) {
  return g ? gcqg::pick_source(g->host, seed, k) : -1;
}

extern "C" int gcq_graph_destroy(gcq_graph* g) {
  if (!g) return GCQ_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  ctx_free(g);
  if (g->d_off) (void)hipFree(g->d_off);
  if (g->d_idx) (void)hipFree(g->d_idx);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return GCQ_OK;
}

extern "C" int gcq_bfs_run(gcq_graph* g, int32_t src, const gcq_bfs_config* cfg, int32_t* depth_out,
                           gcq_bfs_result* res) {
  GCQ_REQUIRE(g && cfg && res, "null argument");
  GCQ_REQUIRE(cfg->block >= 64 && cfg->block % 64 == 0 && cfg->block <= 1024, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->threads >= cfg->block && cfg->threads % cfg->block == 0,
              "threads must be a positive multiple of block (whole waves, bfs_bench.cpp:521)");
  GCQ_REQUIRE(cfg->mode == GCQ_BFS_HOST_LOOP || cfg->mode == GCQ_BFS_PERSISTENT, "bad mode");
  std::memset(res, 0, sizeof *res);
  if (src < 0 || src >= g->host.nv) src = 0;  // bfs_bench.cpp:555
  int rc = ensure_uploaded(g);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  rc = ctx_ensure(g, cfg->kind, cfg->threads);
  if (rc) return rc;
  rc = reset_run(g, src);
  if (rc) return rc;
  BfsCtx& c = g->ctx;
  double kms = 0;
  switch (c.kind) {
    case GCQ_GWFQ: {
      bfsk::Frontiers<bfsk::FGwf> fq;
      for (int i = 0; i < 3; ++i) fq.q[i] = bfsk::FGwf{(gcq::gwf::Ring*)c.q[i % 2]};
      rc = traverse(g, fq, cfg, &kms);
      break;
    }
    case GCQ_GLFQ: {
      bfsk::Frontiers<bfsk::FLfq> fq;
      for (int i = 0; i < 3; ++i) fq.q[i] = bfsk::FLfq{(gcq::lfq::Ring*)c.q[i % 2]};
      rc = traverse(g, fq, cfg, &kms);
      break;
    }
    case GCQ_SFQ: {
      bfsk::Frontiers<bfsk::FSfq> fq;
      for (int i = 0; i < 3; ++i) fq.q[i] = bfsk::FSfq{(gcq::sfq::Queue*)c.q[i % 2]};
      rc = traverse(g, fq, cfg, &kms);
      break;
    }
    case GCQ_WFQ: {
      bfsk::Frontiers<bfsk::FYmc> fq;
      for (int i = 0; i < 3; ++i) fq.q[i] = bfsk::FYmc{(gcq::ymc::Queue*)c.q[i], (gcq::ymc::Handle*)c.h[i]};
      rc = traverse(g, fq, cfg, &kms);
      break;
    }
  }
  if (rc) return rc;
  bfsk::Ctl h;
  GCQ_HIP(hipMemcpyAsync(&h, c.ctl, sizeof h, hipMemcpyDeviceToHost, g->stream));
  if (depth_out)
    GCQ_HIP(hipMemcpyAsync(depth_out, c.level, 4ull * g->host.nv, hipMemcpyDeviceToHost, g->stream));
  GCQ_HIP(hipStreamSynchronize(g->stream));
  res->edges_scanned = h.edges;
  res->reached = h.reached + 1;
  res->levels = h.levels;
  res->kernel_ms = kms;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (h.abort) {
    gcqh::set_error("BFS aborted: a grid-barrier wait exceeded its spin limit");
    return GCQ_ETIMEOUT;
  }
  if (h.err) {
    gcqh::set_error("BFS frontier enqueue failed (queue capacity or cell pool exhausted)");
    return GCQ_EFULL;
  }
  return GCQ_OK;
}
