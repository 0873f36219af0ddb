// core.hip — libgcq: queue lifecycle, atomic-roofline microbenchmarks, the
// timed throughput harness, the FIFO self-check and the logged history runs.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <filesystem>
#include <mutex>
#include <vector>

#include "queue_impl.hpp"

// ======================================================================== errors
namespace gcqh {
static thread_local char g_err[512] = {0};
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
}  // namespace gcqh

extern "C" const char* gcq_last_error(void) { return gcqh::g_err; }
extern "C" const char* gcq_version(void) { return "gcq-mi355x 0.1 (gfx950)"; }
extern "C" const char* gcq_strerror(int rc) {
  switch (rc) {
    case GCQ_OK: return "ok";
    case GCQ_EINVAL: return "invalid argument";
    case GCQ_ENOMEM: return "device allocation failed";
    case GCQ_EHIP: return "HIP runtime error";
    case GCQ_ETIMEOUT: return "spin limit / watchdog expired";
    case GCQ_EFULL: return "capacity exhausted";
    case GCQ_EVERIFY: return "device self-check failed";
    case GCQ_ENODEV: return "no usable GPU";
  }
  return "unknown error";
}

extern "C" int gcq_device_info(char* name, size_t name_len, int* cu_count, uint64_t* hbm_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  hipDeviceProp_t p{};
  GCQ_HIP(hipGetDeviceProperties(&p, 0));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  return GCQ_OK;
}

// ======================================================================== lifecycle
namespace {

__global__ void k_init_gwf(void* base, gcq::gwf::InitParams p) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::gwf::Mpmc m = gcq::gwf::mpmc_view(base, p.n, p.max_threads);
  gcq::gwf::init_slice(m.aq, p, false, gid, st);
  gcq::gwf::init_slice(m.fq, p, true, gid, st);
  for (uint32_t i = gid; i < p.n; i += st) m.data[i] = 0;
}

__global__ void k_init_lfq(void* base, uint32_t n, uint32_t remap_log2, uint32_t poll_backoff) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::lfq::Mpmc m = gcq::lfq::mpmc_view(base, n);
  gcq::lfq::init_slice(m.aq, n, false, remap_log2, gid, st, poll_backoff);
  gcq::lfq::init_slice(m.fq, n, true, remap_log2, gid, st);
  for (uint32_t i = gid; i < n; i += st) m.data[i] = 0;
}

__global__ void k_init_sfq(void* base, uint64_t qlen, uint32_t nprocs) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::sfq::init_slice((gcq::sfq::Queue*)base, qlen, nprocs, gid, st);
}

// Zero cells [0, upto) with a fill kernel (safe under graph capture).
__global__ void k_zero_u64(uint64_t* p, uint64_t words) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < words; i += st) p[i] = 0;
}

__global__ void k_init_ymc(void* base, void* handles, uint64_t cells, uint32_t nthreads, uint32_t patience) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  gcq::ymc::init_slice((gcq::ymc::Queue*)base, (gcq::ymc::Handle*)handles, cells, nthreads, patience, gid, st);
  // library queues reclaim their cells as they run (ring mode)
  gcq::ymc::init_ring_slice((gcq::ymc::Queue*)base, (gcq::ymc::Handle*)handles, cells, nthreads, gid, st);
}

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
uint32_t grid_for(uint64_t work, uint32_t block = 256, uint32_t cap = 4096) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

int launch_init(gcq_queue* q) {
  const gcq_queue_config& c = q->cfg;
  switch (c.kind) {
    case GCQ_GWFQ: {
      gcq::gwf::InitParams p{c.capacity, c.max_threads, c.remap ? 4u : 0u, c.help_delay, c.enq_patience,
                             c.deq_patience, c.poll_backoff};
      hipLaunchKernelGGL(k_init_gwf, dim3(grid_for(std::max(2u * c.capacity, c.max_threads))), dim3(256), 0,
                         q->stream, q->dev, p);
      break;
    }
    case GCQ_GLFQ:
      hipLaunchKernelGGL(k_init_lfq, dim3(grid_for(2u * c.capacity)), dim3(256), 0, q->stream, q->dev,
                         c.capacity, c.remap ? 4u : 0u, c.poll_backoff);
      break;
    case GCQ_SFQ:
      hipLaunchKernelGGL(k_init_sfq, dim3(grid_for(c.capacity)), dim3(256), 0, q->stream, q->dev,
                         (uint64_t)c.capacity, c.max_threads);
      break;
    case GCQ_WFQ: {
      // Zero the whole cell pool: it is a ring, every cell may have been used.
      const uint64_t upto = c.wfq_segment_cells;
      uint64_t* cells = (uint64_t*)((char*)q->dev + gcq::ymc::kHdrBytes);
      hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(upto * 2, 256, 65536)), dim3(256), 0, q->stream, cells, upto * 2);
      hipLaunchKernelGGL(k_init_ymc, dim3(grid_for(c.max_threads)), dim3(256), 0, q->stream, q->dev, q->aux,
                         (uint64_t)c.wfq_segment_cells, c.max_threads, c.enq_patience);
      break;
    }
    default:
      gcqh::set_error("unknown queue kind %d", c.kind);
      return GCQ_EINVAL;
  }
  GCQ_HIP(hipGetLastError());
  GCQ_HIP(hipStreamSynchronize(q->stream));
  return GCQ_OK;
}

}  // namespace

extern "C" void gcq_queue_config_default(int kind, gcq_queue_config* c) {
  std::memset(c, 0, sizeof *c);
  c->kind = kind;
  c->capacity = 1u << 20;
  c->max_threads = 65536;
  c->remap = 0;
  c->help_delay = 64;     // WF_HELP_DELAY   gwf_queue.hpp:296-298
  c->enq_patience = 128;  // WF_ENQ_PATIENCE gwf_queue.hpp:300-302
  c->deq_patience = 256;  // WF_DEQ_PATIENCE gwf_queue.hpp:304-306
  if (kind == GCQ_WFQ) c->enq_patience = 16;  // WF_PATIENCE wfqueue_hip_opt.hpp:29-31
  c->wfq_segment_cells = 0;
  c->poll_backoff = (kind == GCQ_GWFQ || kind == GCQ_GLFQ) ? 2 : 0;
}

extern "C" int gcq_queue_create(const gcq_queue_config* cfg, gcq_queue** out) {
  GCQ_REQUIRE(cfg && out, "null argument");
  *out = nullptr;
  gcq_queue_config c = *cfg;
  GCQ_REQUIRE(c.max_threads >= 1 && c.max_threads <= (1u << 20), "max_threads must be in [1, 2^20]");
  switch (c.kind) {
    case GCQ_GWFQ:
      GCQ_REQUIRE(is_pow2(c.capacity) && c.capacity >= 64 && c.capacity <= (1u << 25),
                  "G-WFQ capacity must be a power of two in [64, 2^25]");
      GCQ_REQUIRE(c.max_threads < (1u << 20), "G-WFQ thread ids must fit the 20-bit owner field");
      GCQ_REQUIRE(c.max_threads <= c.capacity, "G-WFQ requires max_threads <= capacity (wf_t k <= n)");
      GCQ_REQUIRE(c.enq_patience >= 1 && c.deq_patience >= 1 && c.help_delay >= 1, "patience/help must be >= 1");
      break;
    case GCQ_GLFQ:
      GCQ_REQUIRE(is_pow2(c.capacity) && c.capacity >= 64 && c.capacity <= (1u << 30),
                  "G-LFQ capacity must be a power of two in [64, 2^30]");
      break;
    case GCQ_SFQ:
      GCQ_REQUIRE(is_pow2(c.capacity) && c.capacity >= 64, "SFQ capacity must be a power of two >= 64");
      break;
    case GCQ_WFQ:
      // The cell pool is a ring of segments reclaimed as the queue runs.  It
      // must hold the queue's items plus the reclamation lag; by default it
      // holds max(capacity, 32 cells per thread) cells, a power of two: 2^21
      // (32 MiB) at 65,536 threads, where the split harness refuses no
      // enqueue (at 2^20 cells it refused 1-2% as full in some launches).
      if (c.wfq_segment_cells == 0) {
        const uint64_t want = std::max<uint64_t>({(uint64_t)c.capacity, 64ull * c.max_threads, 1ull << 12});
        uint64_t p2 = 1;
        while (p2 < want) p2 <<= 1;
        GCQ_REQUIRE(p2 <= (1ull << 31), "WFQ cell pool too large");
        c.wfq_segment_cells = (uint32_t)p2;
      }
      GCQ_REQUIRE(is_pow2(c.wfq_segment_cells) && c.wfq_segment_cells >= (1u << 12),
                  "WFQ cell pool must be a power of two >= 4096 cells");
      GCQ_REQUIRE(c.enq_patience <= 1024, "WFQ patience too large");
      break;
    default:
      GCQ_REQUIRE(false, "unknown queue kind %d", c.kind);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  gcq_queue* q = new gcq_queue();
  q->cfg = c;
  switch (c.kind) {
    case GCQ_GWFQ: q->bytes = gcq::gwf::mpmc_bytes(c.capacity, c.max_threads); break;
    case GCQ_GLFQ: q->bytes = gcq::lfq::mpmc_bytes(c.capacity); break;
    case GCQ_SFQ: q->bytes = gcq::sfq::queue_bytes(c.capacity); break;
    case GCQ_WFQ: q->bytes = gcq::ymc::queue_bytes(c.wfq_segment_cells); break;
  }
  auto fail = [&](int rc) {
    if (q->dev) (void)hipFree(q->dev);
    if (q->aux) (void)hipFree(q->aux);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
    return rc;
  };
  if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess) {
    gcqh::set_error("hipStreamCreate failed");
    return fail(GCQ_EHIP);
  }
  if (hipMalloc(&q->dev, q->bytes) != hipSuccess) {
    gcqh::set_error("hipMalloc(%llu) for queue failed", (unsigned long long)q->bytes);
    return fail(GCQ_ENOMEM);
  }
  if (c.kind == GCQ_WFQ) {
    // handles, then one reclamation floor per wave
    q->aux_bytes = sizeof(gcq::ymc::Handle) * (uint64_t)c.max_threads + 8ull * gcq::ymc::wave_slots(c.max_threads);
    if (hipMalloc(&q->aux, q->aux_bytes) != hipSuccess) {
      gcqh::set_error("hipMalloc for WFQ handles failed");
      return fail(GCQ_ENOMEM);
    }
  }
  int rc = launch_init(q);
  if (rc) return fail(rc);
  *out = q;
  return GCQ_OK;
}

extern "C" int gcq_queue_reset(gcq_queue* q) {
  GCQ_REQUIRE(q, "null queue");
  return launch_init(q);
}

extern "C" int gcq_queue_destroy(gcq_queue* q) {
  if (!q) return GCQ_OK;
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  if (q->dev) (void)hipFree(q->dev);
  if (q->aux) (void)hipFree(q->aux);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
  return GCQ_OK;
}

extern "C" void* gcq_queue_device_ptr(gcq_queue* q) { return q ? q->dev : nullptr; }
extern "C" uint64_t gcq_queue_bytes(const gcq_queue* q) { return q ? q->bytes : 0; }

// ======================================================================== FAA peak
namespace {

__global__ void k_faa(uint64_t* base, uint32_t stride_words, uint32_t k, uint32_t iters, uint32_t aggregate,
                      uint32_t poll_pct, uint32_t poll_other, uint64_t* sink) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wave = tid / gcq::kWave;
  uint64_t* c = base + (uint64_t)(wave % k) * stride_words;
  uint64_t acc = 0;
  if (wave % 100u < poll_pct) {  // an idle poller: loads only
    const uint64_t* pl = poll_other ? base + (uint64_t)k * stride_words + 16 : c;
    for (uint32_t i = 0; i < iters; ++i) acc += gcq::ld_rlx(pl);
  } else if (aggregate) {
    for (uint32_t i = 0; i < iters; ++i) acc += gcq::wave_faa(c, (uint64_t)1);
  } else {
    for (uint32_t i = 0; i < iters; ++i) acc += gcq::faa_rlx(c, (uint64_t)1);
  }
  if (acc == 0x5eed5eed5eedull) sink[0] = acc;  // keep the FAAs live
}

__global__ void k_fence(uint64_t* buf, uint32_t mode, uint32_t iters) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t* p = buf + tid;
  for (uint32_t i = 0; i < iters; ++i) {
    gcq::st_rlx(p, (uint64_t)i);
    if (mode == 1) gcq::fence_rel();
    else if (mode == 2) gcq::fence_acq();
    else if (mode == 3) gcq::fence_sc();
  }
}

}  // namespace

extern "C" int gcq_faa_peak(const gcq_faa_config* cfg, gcq_faa_result* res) {
  GCQ_REQUIRE(cfg && res, "null argument");
  GCQ_REQUIRE(cfg->threads > 0 && cfg->block > 0 && cfg->block % 64 == 0 && cfg->block <= 1024,
              "threads > 0, block multiple of 64 <= 1024");
  GCQ_REQUIRE(cfg->counters >= 1 && cfg->stride_bytes >= 8 && cfg->stride_bytes % 8 == 0, "bad counters/stride");
  GCQ_REQUIRE(cfg->iters >= 1 && cfg->launches >= 1, "iters/launches >= 1");
  GCQ_REQUIRE(cfg->poll_percent < 100, "poll_percent must be < 100");
  const uint64_t bytes = (uint64_t)cfg->counters * cfg->stride_bytes + 512;
  uint64_t* d = nullptr;
  GCQ_HIP(hipMalloc(&d, bytes));
  GCQ_HIP(hipMemset(d, 0, bytes));
  hipEvent_t a, b;
  GCQ_HIP(hipEventCreate(&a));
  GCQ_HIP(hipEventCreate(&b));
  const uint32_t grid = (cfg->threads + cfg->block - 1) / cfg->block;
  hipLaunchKernelGGL(k_faa, dim3(grid), dim3(cfg->block), 0, 0, d, cfg->stride_bytes / 8, cfg->counters, cfg->iters,
                     cfg->aggregate, cfg->poll_percent, cfg->poll_other_line, d + bytes / 8 - 1);  // warm-up
  GCQ_HIP(hipEventRecord(a, 0));
  for (uint32_t l = 0; l < cfg->launches; ++l)
    hipLaunchKernelGGL(k_faa, dim3(grid), dim3(cfg->block), 0, 0, d, cfg->stride_bytes / 8, cfg->counters,
                       cfg->iters, cfg->aggregate, cfg->poll_percent, cfg->poll_other_line, d + bytes / 8 - 1);
  GCQ_HIP(hipEventRecord(b, 0));
  GCQ_HIP(hipEventSynchronize(b));
  const double ms = gcqh::ms_between(a, b);
  uint64_t faa_waves = 0;
  for (uint64_t w = 0; w < (uint64_t)grid * cfg->block / 64; ++w) faa_waves += (w % 100u) >= cfg->poll_percent;
  const double tickets = (double)faa_waves * 64 * cfg->iters * cfg->launches;
  res->kernel_ms = ms / cfg->launches;
  res->tickets_per_s = tickets / (ms * 1e-3);
  res->atomics_per_s = cfg->aggregate ? res->tickets_per_s / 64.0 : res->tickets_per_s;
  const uint64_t poll_waves = (uint64_t)grid * cfg->block / 64 - faa_waves;
  res->loads_per_s = (double)poll_waves * cfg->iters * cfg->launches / (ms * 1e-3);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  GCQ_HIP(hipFree(d));
  return GCQ_OK;
}

namespace {
// Streaming copy, 16 B per lane per iteration, grid-stride: the measured
// HBM roofline (and the known-size kernel that calibrates HBM byte counters).
__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < n; i += st) dst[i] = src[i];
}
}  // namespace

extern "C" int gcq_hbm_copy(uint64_t bytes, uint32_t iters, double* gbps, double* kernel_ms) {
  GCQ_REQUIRE(bytes >= 4096 && bytes % 16 == 0 && iters >= 1 && gbps, "bad copy size");
  int dev = 0, cus = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  void *a = nullptr, *b = nullptr;
  GCQ_HIP(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    gcqh::set_error("out of device memory for a %llu-byte copy", (unsigned long long)bytes);
    return GCQ_ENOMEM;
  }
  hipStream_t s;
  GCQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const uint64_t n = bytes / 16;
  const dim3 grid((uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)cus * 32)), blk(256);
  GCQ_HIP(hipMemsetAsync(a, 1, bytes, s));
  hipLaunchKernelGGL(k_copy, grid, blk, 0, s, (const uint4*)a, (uint4*)b, n);  // warm
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  GCQ_HIP(hipEventRecord(e0, s));
  for (uint32_t i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, grid, blk, 0, s, (const uint4*)a, (uint4*)b, n);
  GCQ_HIP(hipEventRecord(e1, s));
  GCQ_HIP(hipEventSynchronize(e1));
  GCQ_HIP(hipGetLastError());
  const double ms = gcqh::ms_between(e0, e1) / iters;
  *gbps = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;  // read + write
  if (kernel_ms) *kernel_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  return GCQ_OK;
}

namespace {
// One lane follows a random cycle through `n` words spaced one 128-B line
// apart, each step's address taken from the previous step's result: the
// latency of one dependent device-scope access.  mode 0: device-scope load,
// 1: device-scope fetch-add of 0 (an atomic served on the memory side),
// 2: load through the caches.
__global__ void k_chase(uint64_t* buf, uint32_t steps, uint32_t mode, uint64_t* sink) {
  if (threadIdx.x != 0) return;
  uint64_t i = 0;
  for (uint32_t k = 0; k < steps; ++k) {
    uint64_t* p = buf + 16 * i;
    if (mode == 0) i = gcq::ld_rlx(p);
    else if (mode == 1) i = gcq::faa_rlx(p, (uint64_t)0);
    else i = gcq::ld_cached(p);
  }
  *sink = i;
}
}  // namespace

extern "C" int gcq_round_trip(uint32_t mode, uint32_t steps, double* ns_per_step) {
  GCQ_REQUIRE(ns_per_step && mode <= 2 && steps >= 16, "bad argument");
  constexpr uint32_t n = 1u << 16;  // 8 MiB of lines: each step a line not touched for ~n steps
  std::vector<uint64_t> h(16ull * n, 0);
  std::vector<uint32_t> order(n);
  for (uint32_t k = 0; k < n; ++k) order[k] = k;
  uint64_t x = 0x9E3779B97F4A7C15ull;  // fixed shuffle: one cycle through every line
  for (uint32_t k = n - 1; k > 0; --k) {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    std::swap(order[k], order[x % (k + 1)]);
  }
  for (uint32_t k = 0; k < n; ++k) h[16ull * order[k]] = order[(k + 1) % n];
  uint64_t* d = nullptr;
  GCQ_HIP(hipMalloc(&d, 8ull * h.size() + 64));
  GCQ_HIP(hipMemcpy(d, h.data(), 8ull * h.size(), hipMemcpyHostToDevice));
  hipEvent_t a, b;
  GCQ_HIP(hipEventCreate(&a));
  GCQ_HIP(hipEventCreate(&b));
  uint64_t* sink = d + h.size();
  hipLaunchKernelGGL(k_chase, dim3(1), dim3(64), 0, 0, d, 256u, mode, sink);  // warm
  GCQ_HIP(hipEventRecord(a, 0));
  hipLaunchKernelGGL(k_chase, dim3(1), dim3(64), 0, 0, d, steps, mode, sink);
  GCQ_HIP(hipEventRecord(b, 0));
  GCQ_HIP(hipEventSynchronize(b));
  GCQ_HIP(hipGetLastError());
  *ns_per_step = gcqh::ms_between(a, b) * 1e6 / steps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(d);
  return GCQ_OK;
}

extern "C" int gcq_fence_cost(uint32_t threads, uint32_t block, uint32_t mode, uint32_t iters, double* ns) {
  GCQ_REQUIRE(ns && threads && block && iters && mode <= 3, "bad argument");
  uint64_t* d = nullptr;
  GCQ_HIP(hipMalloc(&d, 8ull * threads + 64));
  hipEvent_t a, b;
  GCQ_HIP(hipEventCreate(&a));
  GCQ_HIP(hipEventCreate(&b));
  const uint32_t grid = (threads + block - 1) / block;
  hipLaunchKernelGGL(k_fence, dim3(grid), dim3(block), 0, 0, d, mode, 8u);
  GCQ_HIP(hipEventRecord(a, 0));
  hipLaunchKernelGGL(k_fence, dim3(grid), dim3(block), 0, 0, d, mode, iters);
  GCQ_HIP(hipEventRecord(b, 0));
  GCQ_HIP(hipEventSynchronize(b));
  *ns = gcqh::ms_between(a, b) * 1e6 / iters;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  GCQ_HIP(hipFree(d));
  return GCQ_OK;
}

// ======================================================================== throughput
namespace {

// One chunk launch of the throughput harness.  Per-thread counts are reduced
// per wave (shuffles) and per block (LDS) into this block's own stats slot;
// the reference instead issues four device-global atomicAdds per thread
// (throughout_bench.cpp:836-839, 878-881).
// kGroupEnq: producers draw Tail tickets per workgroup (group_faa).  A
// template argument rather than a runtime pointer test, so that grids below
// kCombineEnqMinThreads run the per-wave enqueue code unchanged (with the
// test left in, G-LFQ split 25/75 at 1,024 threads measured 13% slower).
template <class Q, bool kGroupEnq>
__global__ void k_tb(Q q, uint32_t threads, uint32_t chunk, int32_t mode, uint32_t pp, uint32_t launch,
                     uint32_t solo, uint64_t* stats) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ gcq::Combine s_cb, s_ct;
  if (threadIdx.x == 0) {
    gcq::combine_init(&s_cb);
    if (kGroupEnq) gcq::combine_init(&s_ct);
  }
  __syncthreads();
  gcq::Combine* const cb = solo || gridDim.x * blockDim.x < gcq::kCombineMinThreads ? nullptr : &s_cb;
  gcq::Combine* const ct = kGroupEnq ? &s_ct : nullptr;
  uint64_t eo = 0, ef = 0, dok = 0, de = 0;
  q.begin(tid);
  if (tid < threads) {
    const uint64_t seq0 = (uint64_t)launch * chunk + 1;
    if (mode == GCQ_TB_SPLIT) {
      const bool prod = ((uint64_t)tid * 100u) / threads < pp;
      for (uint32_t i = 0; i < chunk; ++i) {
        if (prod) {
          const uint64_t v = ((uint64_t)tid << 32) | (uint32_t)(seq0 + i);
          if (q.enq(tid, v, nullptr, ct)) ++eo; else ++ef;
        } else {
          uint64_t o;
          if (q.deq(tid, &o, nullptr, cb)) ++dok; else ++de;
        }
      }
    } else {
      for (uint32_t i = 0; i < chunk; ++i) {
        const uint64_t v = ((uint64_t)tid << 32) | (uint32_t)(seq0 + i);
        if (q.enq(tid, v, nullptr, ct)) ++eo; else ++ef;
        uint64_t o;
        if (q.deq(tid, &o, nullptr, cb)) ++dok; else ++de;
      }
    }
  }
  q.end(tid);
  __shared__ uint64_t red[4][16];
  const uint32_t w = threadIdx.x / gcq::kWave;
  eo = gcq::wave_sum(eo);
  ef = gcq::wave_sum(ef);
  dok = gcq::wave_sum(dok);
  de = gcq::wave_sum(de);
  if (gcq::lane_id() == 0) {
    red[0][w] = eo;
    red[1][w] = ef;
    red[2][w] = dok;
    red[3][w] = de;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    uint64_t s = 0;
    for (uint32_t k = 0; k < blockDim.x / gcq::kWave; ++k) s += red[threadIdx.x][k];
    stats[blockIdx.x * 4 + threadIdx.x] += s;
  }
}

}  // namespace

extern "C" int gcq_tb_run(gcq_queue* q, const gcq_tb_config* cfg, gcq_tb_result* res) {
  GCQ_REQUIRE(q && cfg && res, "null argument");
  GCQ_REQUIRE(cfg->threads >= 64 && cfg->threads <= q->cfg.max_threads, "threads must be in [64, max_threads]");
  GCQ_REQUIRE(cfg->block % 64 == 0 && cfg->block >= 64 && cfg->block <= 1024, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->chunk_ops >= 1, "chunk_ops >= 1");
  std::memset(res, 0, sizeof *res);
  const uint32_t grid = (cfg->threads + cfg->block - 1) / cfg->block;
  uint64_t* stats = nullptr;
  GCQ_HIP(hipMalloc(&stats, 32ull * grid));
  GCQ_HIP(hipMemsetAsync(stats, 0, 32ull * grid, q->stream));
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  uint32_t launch = 0;
  int rc = gcq_dispatch(
      q,
      [&](auto a) -> int {
        using A = decltype(a);
        const bool group_enq = !cfg->solo_claims && grid * cfg->block >= gcq::kCombineEnqMinThreads;
        auto go = [&]() {
          if (group_enq)
            hipLaunchKernelGGL((k_tb<A, true>), dim3(grid), dim3(cfg->block), 0, q->stream, a, cfg->threads,
                               cfg->chunk_ops, cfg->mode, cfg->producer_percent, launch++, cfg->solo_claims, stats);
          else
            hipLaunchKernelGGL((k_tb<A, false>), dim3(grid), dim3(cfg->block), 0, q->stream, a, cfg->threads,
                               cfg->chunk_ops, cfg->mode, cfg->producer_percent, launch++, cfg->solo_claims, stats);
        };
        using clk = std::chrono::steady_clock;
        // Warmup: host-timed like the reference (throughout_bench.cpp:1010-1021).
        auto t0 = clk::now();
        if (cfg->timed_launches > 0) {
          for (uint32_t i = 0; i < cfg->warmup_launches; ++i) go();
          GCQ_HIP(hipStreamSynchronize(q->stream));
        } else {
          do {
            go();
            GCQ_HIP(hipStreamSynchronize(q->stream));
          } while (std::chrono::duration<double, std::milli>(clk::now() - t0).count() < cfg->warmup_ms);
        }
        // Mean duration of one launch, measured in isolation.
        const int probe = 8;
        GCQ_HIP(hipEventRecord(e0, q->stream));
        for (int i = 0; i < probe; ++i) go();
        GCQ_HIP(hipEventRecord(e1, q->stream));
        GCQ_HIP(hipEventSynchronize(e1));
        res->kernel_ms_mean = gcqh::ms_between(e0, e1) / probe;
        uint32_t per_sync = cfg->launches_per_sync;
        if (per_sync == 0) {
          per_sync = (uint32_t)std::max(1.0, std::min(256.0, 5.0 / std::max(res->kernel_ms_mean, 1e-3)));
        }
        // Counts of the warmup and probe launches, so that a caller can
        // account for every item the run put in or took out.
        {
          std::vector<uint64_t> pre(4ull * grid);
          GCQ_HIP(hipMemcpyAsync(pre.data(), stats, 32ull * grid, hipMemcpyDeviceToHost, q->stream));
          GCQ_HIP(hipStreamSynchronize(q->stream));
          for (uint32_t b = 0; b < grid; ++b) {
            res->pre_enq_success += pre[4 * b + 0];
            res->pre_deq_success += pre[4 * b + 2];
          }
        }
        GCQ_HIP(hipMemsetAsync(stats, 0, 32ull * grid, q->stream));
        uint64_t launches = 0;
        if (cfg->timed_launches > 0) {
          // Launch-count mode: exactly timed_launches back-to-back launches.
          GCQ_HIP(hipStreamSynchronize(q->stream));
          GCQ_HIP(hipEventRecord(e0, q->stream));
          for (uint32_t i = 0; i < cfg->timed_launches; ++i) go();
          launches = cfg->timed_launches;
        } else {
          // Timed window: async launches on one stream, host time checked
          // every per_sync launches; events bracket the whole window.
          t0 = clk::now();
          GCQ_HIP(hipEventRecord(e0, q->stream));
          for (;;) {
            for (uint32_t i = 0; i < per_sync; ++i) go();
            launches += per_sync;
            GCQ_HIP(hipStreamSynchronize(q->stream));
            if (std::chrono::duration<double, std::milli>(clk::now() - t0).count() >= cfg->run_ms) break;
          }
        }
        GCQ_HIP(hipEventRecord(e1, q->stream));
        GCQ_HIP(hipEventSynchronize(e1));
        GCQ_HIP(hipGetLastError());
        res->elapsed_ms = gcqh::ms_between(e0, e1);
        res->launches = launches;
        return GCQ_OK;
      },
      /*exact_empty=*/false, /*blocking=*/cfg->mode == GCQ_TB_BALANCED);
  if (rc == GCQ_OK) {
    std::vector<uint64_t> h(4ull * grid);
    if (hipMemcpy(h.data(), stats, 32ull * grid, hipMemcpyDeviceToHost) != hipSuccess) rc = GCQ_EHIP;
    for (uint32_t b = 0; b < grid; ++b) {
      res->enq_success += h[4 * b + 0];
      res->enq_fail += h[4 * b + 1];
      res->deq_success += h[4 * b + 2];
      res->deq_empty += h[4 * b + 3];
    }
    res->succ_mops = (double)(res->enq_success + res->deq_success) / (res->elapsed_ms * 1e-3) / 1e6;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(stats);
  return rc;
}

// ======================================================================== drain
namespace {

// Every lane dequeues until its dequeue answers EMPTY; on a quiescent queue
// the first EMPTY means the queue is empty (each queue's EMPTY answer is
// linearizable, and nothing is enqueued meanwhile).
template <class Q>
__global__ void k_drain(Q q, uint32_t threads, uint64_t* count) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t n = 0;
  q.begin(tid);
  if (tid < threads) {
    uint64_t v;
    while (q.deq(tid, &v)) ++n;
  }
  q.end(tid);
  n = gcq::wave_sum(n);
  if (gcq::lane_id() == 0 && n) gcq::faa_rlx(count, n);
}

}  // namespace

extern "C" int gcq_queue_drain(gcq_queue* q, uint32_t threads, uint64_t* items) {
  GCQ_REQUIRE(q && items, "null argument");
  GCQ_REQUIRE(threads >= 64 && threads % 64 == 0 && threads <= q->cfg.max_threads,
              "threads must be a multiple of 64 in [64, max_threads]");
  uint64_t* d = nullptr;
  GCQ_HIP(hipMalloc(&d, 8));
  GCQ_HIP(hipMemsetAsync(d, 0, 8, q->stream));
  int rc = gcq_dispatch(q, [&](auto a) -> int {
    using A = decltype(a);
    hipLaunchKernelGGL(k_drain<A>, dim3(threads / 64), dim3(64), 0, q->stream, a, threads, d);
    GCQ_HIP(hipGetLastError());
    return GCQ_OK;
  });
  if (rc == GCQ_OK) {
    if (hipMemcpyAsync(items, d, 8, hipMemcpyDeviceToHost, q->stream) != hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess) {
      gcqh::set_error("drain readback failed");
      rc = GCQ_EHIP;
    }
  }
  (void)hipFree(d);
  return rc;
}

// ======================================================================== FIFO check
namespace {

struct FifoDev {
  uint32_t* counts;    // [producers * ops] receipts per token
  uint64_t* stamps;    // [producers * ops][2] logical call / return time of its dequeue
  uint32_t* errs;      // oob, dup, order, zeros, gt1, produce-fail
  uint64_t* consumed;
  uint64_t* clock;     // logical clock shared by every dequeue
};

constexpr uint32_t kFifoSpinLimit = 1u << 22;

// One dequeue attempt of the drain.  Exactly-once is checked with a receipt
// count per token (throughout_bench.cpp:581-583).  Order is checked after the
// drain from logical call/return stamps: the reference's running atomicMax
// of the last sequence seen (throughout_bench.cpp:585-588) also flags two
// correctly ordered dequeues whose atomicMax calls land in the other order.
template <class Q>
__device__ void fifo_take(const Q& q, uint32_t tid, const FifoDev& f, uint32_t producers, uint32_t ops,
                          uint64_t total, uint32_t* streak, gcq::Combine* cb = nullptr) {
  uint64_t v = 0;
  const uint64_t call = gcq::faa_rlx(f.clock, (uint64_t)1);
  if (!q.deq(tid, &v, nullptr, cb)) {
    ++*streak;
    return;
  }
  const uint64_t ret = gcq::faa_rlx(f.clock, (uint64_t)1);
  *streak = 0;
  const uint64_t ticket = gcq::faa_rlx(f.consumed, (uint64_t)1);
  if (ticket >= total) return;
  const uint32_t ptid = (uint32_t)(v >> 32);
  const uint32_t s = (uint32_t)v;
  const uint32_t seq = s ? s - 1u : 0u;
  if (ptid >= producers || seq >= ops || s == 0) {
    gcq::faa_rlx(&f.errs[0], 1u);
    return;
  }
  const uint64_t item = (uint64_t)ptid * ops + seq;
  if (gcq::faa_rlx(&f.counts[item], 1u) >= 1u) {
    gcq::faa_rlx(&f.errs[1], 1u);
    return;
  }
  f.stamps[2 * item] = call;
  f.stamps[2 * item + 1] = ret;
}

template <class Q>
__global__ void k_fifo_produce(Q q, uint32_t producers, uint32_t ops, uint32_t* fail) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  q.begin(tid);
  if (tid < producers) {
    for (uint32_t i = 0; i < ops; ++i) {
      const uint64_t tok = ((uint64_t)tid << 32) | (uint64_t)(i + 1u);
      uint32_t spins = 0;
      while (!q.enq(tid, tok)) {
        if (++spins > kFifoSpinLimit) {
          gcq::faa_rlx(fail, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (spins > kFifoSpinLimit) break;
    }
  }
  q.end(tid);
}

template <class Q>
__global__ void k_fifo_drain(Q q, uint32_t threads, FifoDev f, uint32_t producers, uint32_t ops, uint64_t total) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  q.begin(tid);
  uint32_t streak = 0;
  while (tid < threads && gcq::ld_rlx(f.consumed) < total && streak < kFifoSpinLimit) {
    fifo_take(q, tid, f, producers, ops, total, &streak);
    if (streak) __builtin_amdgcn_s_sleep(2);
  }
  q.end(tid);
}

// Producers (tid*100/threads < 25) and consumers in one launch.
template <class Q>
__global__ void k_fifo_mixed(Q q, uint32_t threads, uint32_t producers, FifoDev f, uint32_t ops, uint64_t total,
                             uint32_t* fail) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ gcq::Combine s_cb, s_ct;
  if (threadIdx.x == 0) {
    gcq::combine_init(&s_cb);
    gcq::combine_init(&s_ct);
  }
  __syncthreads();
  if (tid >= threads) return;
  q.begin(tid);
  if (tid < producers) {
    for (uint32_t i = 0; i < ops; ++i) {
      const uint64_t tok = ((uint64_t)tid << 32) | (uint64_t)(i + 1u);
      uint32_t spins = 0;
      while (!q.enq(tid, tok, nullptr, &s_ct)) {
        if (++spins > kFifoSpinLimit) {
          gcq::faa_rlx(fail, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (spins > kFifoSpinLimit) break;
    }
    q.end(tid);
    return;
  }
  uint32_t streak = 0;
  while (gcq::ld_rlx(f.consumed) < total && streak < kFifoSpinLimit) {
    fifo_take(q, tid, f, producers, ops, total, &streak, &s_cb);
    if (streak) __builtin_amdgcn_s_sleep(2);
  }
  q.end(tid);
}

__global__ void k_fifo_verify(const uint32_t* counts, uint64_t n, uint32_t* zeros, uint32_t* gt1) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
  uint32_t z = 0, g = 0;
  for (uint64_t i = gid; i < n; i += st) {
    const uint32_t c = counts[i];
    z += c == 0u;
    g += c > 1u;
  }
  if (z) gcq::faa_rlx(zeros, z);
  if (g) gcq::faa_rlx(gt1, g);
}

// Per producer: its tokens were enqueued one after another, so the dequeue
// of token s must not finish before the dequeue of an earlier token starts.
__global__ void k_fifo_order(const uint32_t* counts, const uint64_t* stamps, uint32_t producers, uint32_t ops,
                             uint32_t* order_errs) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= producers) return;
  uint64_t max_call = 0;
  uint32_t bad = 0;
  for (uint32_t s = 0; s < ops; ++s) {
    const uint64_t item = (uint64_t)p * ops + s;
    if (counts[item] != 1u) continue;
    const uint64_t call = stamps[2 * item], ret = stamps[2 * item + 1];
    if (ret < max_call) ++bad;
    if (call > max_call) max_call = call;
  }
  if (bad) gcq::faa_rlx(order_errs, bad);
}

}  // namespace

extern "C" int gcq_fifo_check(gcq_queue* q, uint32_t threads, uint32_t ops, uint32_t block, int32_t concurrent,
                              gcq_fifo_result* res) {
  GCQ_REQUIRE(q && res, "null argument");
  GCQ_REQUIRE(threads >= 1 && threads <= q->cfg.max_threads && ops >= 1, "bad threads/ops");
  GCQ_REQUIRE(block % 64 == 0 && block >= 64 && block <= 1024, "block must be a multiple of 64");
  std::memset(res, 0, sizeof *res);
  const uint32_t producers = concurrent ? (uint32_t)(((uint64_t)threads * 25 + 99) / 100) : threads;
  GCQ_REQUIRE(producers >= 1, "no producers");
  const uint64_t total = (uint64_t)producers * ops;
  if (!concurrent)
    GCQ_REQUIRE(total <= (q->cfg.kind == GCQ_WFQ ? q->cfg.wfq_segment_cells : q->cfg.capacity),
                "produce-then-drain needs threads*ops <= capacity (%llu items)", (unsigned long long)total);
  res->in_count = total;
  const uint64_t bytes = 16ull * total + 4ull * total + 256;
  char* buf = nullptr;
  GCQ_HIP(hipMalloc(&buf, bytes));
  GCQ_HIP(hipMemsetAsync(buf, 0, bytes, q->stream));
  FifoDev f;
  f.stamps = (uint64_t*)buf;
  f.counts = (uint32_t*)(buf + 16ull * total);
  f.errs = (uint32_t*)(buf + 20ull * total);
  f.consumed = (uint64_t*)(buf + 20ull * total + 128);
  f.clock = f.consumed + 8;
  const uint32_t grid = (threads + block - 1) / block;
  int rc = gcq_dispatch(q, [&](auto a) -> int {
    using A = decltype(a);
    if (concurrent) {
      hipLaunchKernelGGL(k_fifo_mixed<A>, dim3(grid), dim3(block), 0, q->stream, a, threads, producers, f, ops, total,
                         f.errs + 5);
    } else {
      hipLaunchKernelGGL(k_fifo_produce<A>, dim3(grid), dim3(block), 0, q->stream, a, producers, ops, f.errs + 5);
      hipLaunchKernelGGL(k_fifo_drain<A>, dim3(grid), dim3(block), 0, q->stream, a, threads, f, producers, ops, total);
    }
    GCQ_HIP(hipGetLastError());
    return GCQ_OK;
  });
  if (rc == GCQ_OK) {
    hipLaunchKernelGGL(k_fifo_verify, dim3(grid_for(total, 256, 65535)), dim3(256), 0, q->stream, f.counts, total,
                       f.errs + 3, f.errs + 4);
    hipLaunchKernelGGL(k_fifo_order, dim3(grid_for(producers, 256, 65535)), dim3(256), 0, q->stream, f.counts,
                       f.stamps, producers, ops, f.errs + 2);
    uint32_t errs[6];
    uint64_t consumed = 0;
    if (hipMemcpyAsync(errs, f.errs, sizeof errs, hipMemcpyDeviceToHost, q->stream) != hipSuccess ||
        hipMemcpyAsync(&consumed, f.consumed, 8, hipMemcpyDeviceToHost, q->stream) != hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess) {
      gcqh::set_error("FIFO check readback failed");
      rc = GCQ_EHIP;
    } else {
      res->out_count = std::min<uint64_t>(consumed, total);
      res->err_oob = errs[0];
      res->err_dup = errs[1];
      res->err_order = errs[2];
      res->err_missing = errs[3];
      res->err_gt1 = errs[4];
      res->verified = res->out_count == res->in_count && !errs[0] && !errs[1] && !errs[2] && !errs[3] && !errs[4] &&
                      !errs[5];
    }
  }
  (void)hipFree(buf);
  return rc;
}

// ======================================================================== history
namespace {

// Counts shared-memory steps of one operation and, for chosen threads,
// stalls between taking a ticket and touching its slot (the point where a
// preempted thread hurts a lock-free queue most).
struct StepProbe {
  uint32_t steps = 0, slow = 0;
  uint32_t delay_cycles = 0;
  __device__ __forceinline__ void step() { ++steps; }
  __device__ __forceinline__ void after_ticket() {
    if (delay_cycles) gcq::spin_cycles(delay_cycles);
  }
  __device__ __forceinline__ void slow_entry() { slow = 1; }
};

struct LogEntry {  // device-side mirror of gcq_hist_entry
  uint32_t proc, op;
  uint64_t arg, ret, call, ret_time;
  uint32_t steps, slow;
};
static_assert(sizeof(LogEntry) == sizeof(gcq_hist_entry), "log entry layout");

template <class Q>
__global__ void k_hist(Q q, gcq_hist_config c, uint32_t stride, LogEntry* log, uint64_t* clock) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ gcq::Combine s_cb, s_ct;
  if (threadIdx.x == 0) {
    gcq::combine_init(&s_cb);
    gcq::combine_init(&s_ct);
  }
  __syncthreads();
  if (tid >= c.threads) return;
  q.begin(tid);
  const bool victim = c.delay_every && ((tid + (uint32_t)c.seed) % c.delay_every == 0);
  const bool prod = ((uint64_t)tid * 100u) / c.threads < c.producer_percent;
  LogEntry* my = log + (uint64_t)tid * stride;
  uint32_t k = 0;
  for (uint32_t i = 0; i < c.ops; ++i) {
    const bool do_enq = c.mode == 0 || prod;
    const bool do_deq = c.mode == 0 || !prod;
    if (do_enq) {
      StepProbe p;
      p.delay_cycles = victim ? c.delay_cycles : 0;
      const uint64_t v = ((uint64_t)tid << 32) | (uint64_t)(i + 1u);
      LogEntry e;
      e.proc = tid;
      e.op = 0;
      e.arg = v;
      e.call = gcq::faa_rlx(clock, (uint64_t)1);
      const bool ok = q.enq(tid, v, &p, &s_ct);
      e.ret_time = gcq::faa_rlx(clock, (uint64_t)1);
      e.ret = ok ? 1 : 0;
      e.steps = p.steps;
      e.slow = p.slow;
      my[k++] = e;
    }
    if (do_deq) {
      StepProbe p;
      p.delay_cycles = victim ? c.delay_cycles : 0;
      LogEntry e;
      e.proc = tid;
      e.op = 1;
      e.arg = 0;
      e.call = gcq::faa_rlx(clock, (uint64_t)1);
      uint64_t out = 0;
      const bool ok = q.deq(tid, &out, &p, &s_cb);
      e.ret_time = gcq::faa_rlx(clock, (uint64_t)1);
      e.ret = ok ? out : 0;
      e.steps = p.steps;
      e.slow = p.slow;
      my[k++] = e;
    }
  }
  q.end(tid);
}

}  // namespace

extern "C" int gcq_hist_run(gcq_queue* q, const gcq_hist_config* cfg, gcq_hist_entry* out, size_t cap,
                            size_t* n_out) {
  GCQ_REQUIRE(q && cfg && out && n_out, "null argument");
  GCQ_REQUIRE(cfg->threads >= 1 && cfg->threads <= q->cfg.max_threads && cfg->ops >= 1, "bad threads/ops");
  GCQ_REQUIRE(cfg->block % 64 == 0 && cfg->block >= 64 && cfg->block <= 1024, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->mode <= 1, "mode must be 0 or 1");
  const uint32_t stride = (cfg->mode == 0 ? 2u : 1u) * cfg->ops;
  const uint64_t slots = (uint64_t)cfg->threads * stride;
  LogEntry* log = nullptr;
  uint64_t* clock = nullptr;
  GCQ_HIP(hipMalloc(&log, sizeof(LogEntry) * slots));
  GCQ_HIP(hipMalloc(&clock, 256));
  GCQ_HIP(hipMemsetAsync(log, 0, sizeof(LogEntry) * slots, q->stream));
  const uint64_t one = 1;
  GCQ_HIP(hipMemcpyAsync(clock, &one, 8, hipMemcpyHostToDevice, q->stream));
  const uint32_t grid = (cfg->threads + cfg->block - 1) / cfg->block;
  int rc = gcq_dispatch(q, [&](auto a) -> int {
    using A = decltype(a);
    hipLaunchKernelGGL(k_hist<A>, dim3(grid), dim3(cfg->block), 0, q->stream, a, *cfg, stride, log, clock);
    GCQ_HIP(hipGetLastError());
    return GCQ_OK;
  });
  if (rc == GCQ_OK) {
    std::vector<LogEntry> h(slots);
    if (hipMemcpyAsync(h.data(), log, sizeof(LogEntry) * slots, hipMemcpyDeviceToHost, q->stream) != hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess) {
      gcqh::set_error("history readback failed: %s", hipGetErrorString(hipGetLastError()));
      rc = GCQ_EHIP;
    } else {
      std::vector<LogEntry> used;
      used.reserve(slots);
      for (auto& e : h)
        if (e.ret_time != 0) used.push_back(e);
      std::sort(used.begin(), used.end(), [](const LogEntry& x, const LogEntry& y) {
        if (x.call != y.call) return x.call < y.call;
        if (x.ret_time != y.ret_time) return x.ret_time < y.ret_time;
        if (x.proc != y.proc) return x.proc < y.proc;
        return x.op < y.op;
      });
      const size_t n = std::min(cap, used.size());
      std::memcpy(out, used.data(), n * sizeof(LogEntry));
      *n_out = used.size();
      if (used.size() > cap) {
        gcqh::set_error("history has %zu entries, buffer holds %zu", used.size(), cap);
        rc = GCQ_EINVAL;
      }
    }
  }
  (void)hipFree(log);
  (void)hipFree(clock);
  return rc;
}

extern "C" int gcq_hist_write_jsonl(const gcq_hist_entry* e, size_t n, const char* path) {
  GCQ_REQUIRE(e || n == 0, "null entries");
  GCQ_REQUIRE(path, "null path");
  // The parent directory is created as the reference's dump does
  // (history.cpp:249-255).
  const std::filesystem::path out_path(path);
  if (out_path.has_parent_path()) {
    std::error_code ec;
    std::filesystem::create_directories(out_path.parent_path(), ec);
    GCQ_REQUIRE(!ec, "cannot create the directory of %s", path);
  }
  FILE* f = fopen(path, "wb");
  GCQ_REQUIRE(f, "cannot open %s", path);
  for (size_t i = 0; i < n; ++i)
    fprintf(f, "{\"proc\":%u,\"op\":%u,\"arg\":%llu,\"ret\":%llu,\"call\":%llu,\"end\":%llu}\n", e[i].proc, e[i].op,
            (unsigned long long)e[i].arg, (unsigned long long)e[i].ret, (unsigned long long)e[i].call,
            (unsigned long long)e[i].ret_time);
  fclose(f);
  return GCQ_OK;
}
