// rt_compact.hip — wavefront ray tracing with stream compaction, the
// comparator of the queue renderer (rt.hip).
//
// Algorithm of the reference's compaction baseline
// (concurrent_rt/baselines/rt_compaction.cpp:219-472, 748-833): trace every
// live ray of a bounce, mark the rays that continue, pack them contiguously
// for the next bounce, repeat until none is left or the depth limit is
// reached.  The shading is rt_final.cpp's (rt_shade.hpp), not the
// baseline's own weights, so this renderer and the queue renderer compute
// the same image and differ only in how they keep the live rays packed:
// a prefix sum over continuation flags here, per-tile queues there.
//
// MI355X design (own):
//  * one launch per bounce, the compaction fused into the trace kernel as a
//    single-pass chained scan (decoupled look-back, Merrill & Garland 2016):
//    a block counts its continuing rays with wave ballots, publishes the
//    count, adds up its predecessors' published counts with one wave-wide
//    look-back, and writes its continuing rays to [prefix, prefix + count)
//    of the next bounce's buffer — no flag array, no separate scan or
//    scatter launches (the baseline uses four launches and a host-side scan
//    of block sums per bounce);
//  * block ids come from an atomic ticket in launch order, so a block only
//    ever waits on blocks that are already running;
//  * primary rays are generated in registers by the first bounce (no
//    generate launch, no primary-ray buffer);
//  * a pixel's running sum rides in its ray record and is stored once, by
//    the bounce that ends the chain, so the additions happen in bounce order
//    and the image equals the queue renderer's bit for bit;
//  * the host reads the next bounce's ray count (4 bytes) between bounces,
//    as the baseline does.
#pragma clang fp contract(off)
#include <algorithm>
#include <chrono>
#include <cstring>

#include "frontier.hpp"
#include "rt_handle.hpp"

namespace rtk {
namespace {

// Scan status word of a block: flag in the top two bits, count below.
constexpr uint64_t kStAgg = 1ull << 62;  // own count published
constexpr uint64_t kStInc = 2ull << 62;  // inclusive prefix published
constexpr uint64_t kStMask = 3ull << 62;
constexpr uint32_t kSpinLimit = 1u << 20;  // look-back rounds before giving up (seconds)

struct CJob {
  int32_t W, H, x0, y0, w, h, max_depth, first;
  uint32_t n_in;        // rays of this bounce
  const RayRec* in;     // packed records (unused by the first bounce)
  RayRec* out;          // the next bounce's records
  float* img;           // region w*h RGB
  uint64_t* status;     // per block
  uint32_t* ctl;        // [0] block ticket, [1] next bounce's count, [2] error
};

// Records between launches: written and read by different kernels, so plain
// 16-byte vector stores and loads; the pixel index rides in the pad word.
__device__ __forceinline__ void put(RayRec* p, const Ray& r, uint32_t pix) {
  float4* q = (float4*)p;
  q[0] = make_float4(r.o.x, r.o.y, r.o.z, __int_as_float(r.depth));
  q[1] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(pix));
  q[2] = make_float4(r.at.x, r.at.y, r.at.z, 0.f);
  q[3] = make_float4(r.acc.x, r.acc.y, r.acc.z, 0.f);
}
__device__ __forceinline__ Ray get(const RayRec* p, uint32_t* pix) {
  const float4* q = (const float4*)p;
  const float4 a = q[0], b = q[1], c = q[2], d = q[3];
  Ray r;
  r.o = v3(a.x, a.y, a.z);
  r.depth = __float_as_int(a.w);
  r.d = v3(b.x, b.y, b.z);
  *pix = __float_as_uint(b.w);
  r.at = v3(c.x, c.y, c.z);
  r.acc = v3(d.x, d.y, d.z);
  return r;
}

__global__ void __launch_bounds__(1024) k_bounce(Scene sc, CJob jb) {
  __shared__ uint32_t s_bid, s_prefix;
  __shared__ uint32_t s_wcount[16];
  const uint32_t nw = blockDim.x / gcq::kWave, w = threadIdx.x / gcq::kWave;
  if (threadIdx.x == 0) s_bid = gcq::faa_rlx(&jb.ctl[0], 1u);
  __syncthreads();
  const uint32_t bid = s_bid;
  const uint32_t i = bid * blockDim.x + threadIdx.x;
  bool cont = false;
  Ray r;
  uint32_t pix = 0;
  if (i < jb.n_in) {
    if (jb.first) {
      pix = i;
      r = primary(sc, jb.W, jb.H, jb.x0 + (int32_t)(i % (uint32_t)jb.w), jb.y0 + (int32_t)(i / (uint32_t)jb.w));
    } else {
      r = get(&jb.in[i], &pix);
    }
    cont = shade(sc, jb.max_depth, norm(sc.light_raw), r);
    if (!cont) {
      jb.img[3 * pix + 0] = r.acc.x;
      jb.img[3 * pix + 1] = r.acc.y;
      jb.img[3 * pix + 2] = r.acc.z;
    }
  }
  const uint64_t ball = __ballot(cont);
  if (gcq::lane_id() == 0) s_wcount[w] = (uint32_t)__builtin_popcountll(ball);
  __syncthreads();
  if (w == 0) {
    // Wave 0: block total, then the look-back over up to 64 predecessors at
    // a time (lane k reads block bid-1-k).
    uint32_t agg = 0;
    for (uint32_t k = 0; k < nw; ++k) agg += s_wcount[k];
    if (gcq::lane_id() == 0) gcq::st_rlx(&jb.status[bid], (bid == 0 ? kStInc : kStAgg) | agg);
    uint32_t prefix = 0;
    if (bid > 0) {
      int64_t j = (int64_t)bid - 1;
      uint32_t spins = 0;
      for (;;) {
        const int64_t mine = j - (int64_t)gcq::lane_id();
        uint64_t st = mine >= 0 ? gcq::ld_rlx(&jb.status[mine]) : kStInc;  // before block 0: prefix 0
        const uint64_t empty = __ballot((st & kStMask) == 0);
        const uint64_t inc = __ballot((st & kStMask) == kStInc);
        // Lanes up to the nearest inclusive one must all be published.
        const uint32_t stop = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
        const uint64_t need = stop >= 63 ? ~0ull : ((2ull << stop) - 1ull);
        if (empty & need) {
          if (++spins > kSpinLimit) {
            if (gcq::lane_id() == 0) gcq::or_rlx(&jb.ctl[2], 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const uint32_t v = gcq::lane_id() <= stop ? (uint32_t)(st & 0xFFFFFFFFull) : 0u;
        prefix += (uint32_t)gcq::wave_sum((uint64_t)v);
        if (inc) break;
        j -= 64;
      }
      if (gcq::lane_id() == 0) gcq::st_rlx(&jb.status[bid], kStInc | (uint64_t)(prefix + agg));
    }
    if (gcq::lane_id() == 0) {
      s_prefix = prefix;
      if ((bid + 1) * blockDim.x >= jb.n_in) gcq::st_rlx(&jb.ctl[1], prefix + agg);  // last block
    }
  }
  __syncthreads();
  if (cont) {
    uint32_t off = s_prefix;
    for (uint32_t k = 0; k < w; ++k) off += s_wcount[k];
    put(&jb.out[off + gcq::rank_in(ball)], r, pix);
  }
}

}  // namespace

void free_compact(gcq_rt_scene* s) {
  for (void* p : {(void*)s->cbuf[0], (void*)s->cbuf[1], (void*)s->cstatus, (void*)s->cctl, (void*)s->cimg})
    if (p) (void)hipFree(p);
  s->cbuf[0] = s->cbuf[1] = nullptr;
  s->cstatus = nullptr;
  s->cctl = nullptr;
  s->cimg = nullptr;
  s->cbuf_cap = s->cimg_cap = 0;
  s->cstatus_cap = 0;
}

}  // namespace rtk

extern "C" int gcq_rt_render_compact(gcq_rt_scene* s, const gcq_rt_config* cfg, float* img_out, gcq_rt_result* res) {
  GCQ_REQUIRE(s && cfg && res, "null argument");
  GCQ_REQUIRE(cfg->block >= 64 && cfg->block <= 1024 && cfg->block % 64 == 0, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->width > 0 && cfg->height > 0 && cfg->max_depth >= 0, "bad image geometry");
  const int32_t w = cfg->w ? cfg->w : cfg->width, h = cfg->h ? cfg->h : cfg->height;
  GCQ_REQUIRE(cfg->x0 >= 0 && cfg->y0 >= 0 && w > 0 && h > 0 && cfg->x0 + w <= cfg->width && cfg->y0 + h <= cfg->height,
              "region outside the image");
  const uint64_t npx = (uint64_t)w * h;
  GCQ_REQUIRE(npx < (1ull << 31), "region too large");
  std::memset(res, 0, sizeof *res);
  int rc = rtk::scene_upload(s);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t B = cfg->block;
  const uint32_t max_blocks = (uint32_t)((npx + B - 1) / B);
  if (npx > s->cbuf_cap) {
    for (int k = 0; k < 2; ++k) {
      if (s->cbuf[k]) (void)hipFree(s->cbuf[k]);
      s->cbuf[k] = nullptr;
    }
    s->cbuf_cap = 0;
    GCQ_HIP(hipMalloc(&s->cbuf[0], sizeof(rtk::RayRec) * npx));
    GCQ_HIP(hipMalloc(&s->cbuf[1], sizeof(rtk::RayRec) * npx));
    s->cbuf_cap = npx;
  }
  if (npx > s->cimg_cap) {
    if (s->cimg) (void)hipFree(s->cimg);
    s->cimg = nullptr;
    s->cimg_cap = 0;
    GCQ_HIP(hipMalloc(&s->cimg, 12ull * npx));
    s->cimg_cap = npx;
  }
  if (max_blocks > s->cstatus_cap) {
    if (s->cstatus) (void)hipFree(s->cstatus);
    s->cstatus = nullptr;
    s->cstatus_cap = 0;
    GCQ_HIP(hipMalloc(&s->cstatus, 8ull * max_blocks));
    s->cstatus_cap = max_blocks;
  }
  if (!s->cctl) GCQ_HIP(hipMalloc(&s->cctl, 64));

  const rtk::Scene sc = rtk::scene_view(s);
  rtk::CJob jb;
  jb.W = cfg->width;
  jb.H = cfg->height;
  jb.x0 = cfg->x0;
  jb.y0 = cfg->y0;
  jb.w = w;
  jb.h = h;
  jb.max_depth = cfg->max_depth;
  jb.img = s->cimg;
  jb.status = s->cstatus;
  jb.ctl = s->cctl;
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  GCQ_HIP(hipEventRecord(e0, s->stream));
  uint32_t n = (uint32_t)npx, cur = 0, err = 0;
  uint64_t total = 0;
  for (int32_t depth = 0; depth <= cfg->max_depth && n > 0; ++depth) {
    const uint32_t nb = (n + B - 1) / B;
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3((nb + 255) / 256), dim3(256), 0, s->stream, s->cstatus, (uint64_t)nb);
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(64), 0, s->stream, (uint64_t*)s->cctl, (uint64_t)8);
    jb.first = depth == 0;
    jb.n_in = n;
    jb.in = s->cbuf[cur];
    jb.out = s->cbuf[cur ^ 1];
    hipLaunchKernelGGL(rtk::k_bounce, dim3(nb), dim3(B), 0, s->stream, sc, jb);
    GCQ_HIP(hipGetLastError());
    total += n;
    uint32_t c[3] = {0, 0, 0};
    GCQ_HIP(hipMemcpyAsync(c, s->cctl, sizeof c, hipMemcpyDeviceToHost, s->stream));
    GCQ_HIP(hipStreamSynchronize(s->stream));
    err |= c[2];
    // At the depth limit shade() ends every chain, so the count is 0 there.
    n = c[1];
    cur ^= 1;
  }
  GCQ_HIP(hipEventRecord(e1, s->stream));
  if (img_out) GCQ_HIP(hipMemcpyAsync(img_out, s->cimg, 12ull * npx, hipMemcpyDeviceToHost, s->stream));
  GCQ_HIP(hipStreamSynchronize(s->stream));
  res->kernel_ms = gcqh::ms_between(e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->primary_rays = npx;
  res->total_rays = total;
  res->tiles = 0;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (err) {
    gcqh::set_error("compaction renderer: a block's look-back passed the spin limit");
    return GCQ_ETIMEOUT;
  }
  if (n != 0) {
    gcqh::set_error("compaction renderer: %u rays left after the depth limit", n);
    return GCQ_EVERIFY;
  }
  return GCQ_OK;
}
