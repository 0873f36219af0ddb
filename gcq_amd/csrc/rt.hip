// rt.hip — queue-compacted Whitted ray tracer (the concurrent_rt workload).
//
// Image semantics are those of rt_final.cpp: camera rays (:482-495),
// nearest hit over spheres then planes (:368-419), direct light + ambient,
// mirror bounces that carry (1 - refl) of the hit's colour to the pixel and
// spawn a reflected ray with attenuation albedo * refl (:583-620), sky on a
// miss (:622-627), every contribution clamped to [0, 1] and summed per pixel
// in bounce order (:424-429).
//
// MI355X design (own):
//  * one launch per render: each tile's workers first trace their share of
//    primary rays themselves (no queue round trip for a fresh ray), then
//    drain the tile's queue, which holds the local pixel index of every
//    live bounce ray — the queue compacts the surviving rays onto whatever
//    lanes are free;
//  * a bounce ray travels as a 64-byte record (origin, direction, depth,
//    attenuation and the pixel's running sum) written with device-scope
//    stores that complete (s_waitcnt) before the index is enqueued — no
//    __threadfence (a full-L2 writeback on gfx950);
//  * a pixel's sum lives in its ray record and is stored to the image once,
//    by the lane that ends the chain: no float atomics, and the additions
//    happen in bounce order, as in the reference;
//  * per-tile live-ray counters start at the tile's pixel count and are
//    decremented once per wave step; a wave only reads its counter when it
//    has no work;
//  * floating-point contraction is off in this file so the arithmetic is the
//    plain IEEE sequence of the source (the CPU restatement compiles the
//    same way).
#pragma clang fp contract(off)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "rt_handle.hpp"

namespace rtk {

struct Job {
  int32_t W, H, x0, y0, w, h, tile_w, tile_h, tiles_x, max_depth;
  uint32_t threads_per_tile;
  RayRec* recs;        // region pixel -> record
  float* img;          // region w*h RGB
  uint64_t* alive;     // per tile, 16 words apart (128 B)
  uint64_t* traced;    // per tile, 16 words apart
  uint32_t* err;       // bit 0: enqueue failed, bit 1: idle spin limit
};
constexpr uint32_t kCounterStride = 16;
constexpr uint32_t kIdleLimit = 1u << 24;

// All tiles of a region in one launch: blockIdx.y is the tile, the
// threads_per_tile threads of blockIdx.x are its workers.
template <class F>
__global__ void k_render(const F* __restrict__ queues, Scene sc, Job jb) {
  const uint32_t tile = blockIdx.y;
  const uint32_t wt = blockIdx.x * blockDim.x + threadIdx.x;  // worker id within the tile
  const uint32_t T = jb.threads_per_tile;
  const F q = queues[tile];
  const int32_t tx = (int32_t)(tile % (uint32_t)jb.tiles_x), ty = (int32_t)(tile / (uint32_t)jb.tiles_x);
  const int32_t lx0 = tx * jb.tile_w, ly0 = ty * jb.tile_h;  // tile origin inside the region
  const int32_t tw = min(jb.tile_w, jb.w - lx0), th = min(jb.tile_h, jb.h - ly0);
  const uint32_t npx = (uint32_t)(tw * th);
  const V3 light = norm(sc.light_raw);
  uint64_t* alive = &jb.alive[tile * kCounterStride];
  uint32_t next_primary = wt;
  uint64_t traced = 0;
  uint32_t idle = 0;
  bool held = false;  // this lane's ray continues here (its primaries are done)
  uint32_t i = 0;
  Ray r;
  for (;;) {
    bool have = held;
    if (!held) {
      if (next_primary < npx) {
        i = next_primary;
        next_primary += T;
        r = primary(sc, jb.W, jb.H, jb.x0 + lx0 + (int32_t)(i % (uint32_t)tw),
                    jb.y0 + ly0 + (int32_t)(i / (uint32_t)tw));
        have = true;
      } else if (q.pop(wt, &i)) {
        r = load_ray(&jb.recs[(ly0 + (int32_t)(i / (uint32_t)tw)) * jb.w + lx0 + (int32_t)(i % (uint32_t)tw)]);
        have = true;
      }
    }
    if (__ballot(have) == 0) {
      // No primaries left for this wave and the queue looked empty: done
      // once the tile has no live ray.
      uint64_t a = 0;
      if (gcq::lane_id() == 0) a = gcq::ld_rlx(alive);
      a = gcq::bcast64(a, 0);
      if (a == 0) break;
      if (++idle > kIdleLimit) {
        if (gcq::lane_id() == 0) gcq::or_rlx(jb.err, 2u);
        break;
      }
      __builtin_amdgcn_s_sleep(4);
      continue;
    }
    idle = 0;
    uint32_t ended = 0;
    if (have) {
      ++traced;
      const uint32_t rp = (uint32_t)((ly0 + (int32_t)(i / (uint32_t)tw)) * jb.w + lx0 + (int32_t)(i % (uint32_t)tw));
      if (shade(sc, jb.max_depth, light, r)) {
        // While the lane still has primaries, the bounce ray goes through
        // the tile's queue to whichever lane is free; once they are done
        // the lane traces it on itself, without the queue's round trip.
        held = next_primary >= npx;
        if (!held) {
          store_ray(&jb.recs[rp], r);
          gcq::vm_wait();  // record complete before its index is published
          if (!q.push(wt, i)) {
            gcq::or_rlx(jb.err, 1u);
            ended = 1;
          }
        }
      } else {
        jb.img[3 * rp + 0] = r.acc.x;
        jb.img[3 * rp + 1] = r.acc.y;
        jb.img[3 * rp + 2] = r.acc.z;
        ended = 1;
        held = false;
      }
    }
    const uint32_t e = gcq::wave_sum((uint64_t)ended);
    if (gcq::lane_id() == 0 && e) gcq::faa_rlx(alive, (uint64_t)0 - (uint64_t)e);
  }
  const uint64_t tr = gcq::wave_sum(traced);
  if (gcq::lane_id() == 0 && tr) gcq::faa_rlx(&jb.traced[tile * kCounterStride], tr);
}

__global__ void k_init_alive(uint64_t* alive, uint64_t* traced, Job jb, uint32_t ntiles) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int32_t tx = (int32_t)(t % (uint32_t)jb.tiles_x), ty = (int32_t)(t / (uint32_t)jb.tiles_x);
  const int32_t tw = min(jb.tile_w, jb.w - tx * jb.tile_w), th = min(jb.tile_h, jb.h - ty * jb.tile_h);
  alive[t * kCounterStride] = (uint64_t)(tw * th);
  traced[t * kCounterStride] = 0;
}

// Per-tile queues of one kind, laid out back to back.
__global__ void k_init_tiles_lfq(char* pool, uint64_t stride, uint32_t n, uint32_t ntiles) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (uint32_t t = 0; t < ntiles; ++t) gcq::lfq::init_slice((gcq::lfq::Ring*)(pool + t * stride), n, false, 0, gid, st, 0);
}
__global__ void k_init_tiles_gwf(char* pool, uint64_t stride, gcq::gwf::InitParams p, uint32_t ntiles) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (uint32_t t = 0; t < ntiles; ++t) gcq::gwf::init_slice((gcq::gwf::Ring*)(pool + t * stride), p, false, gid, st);
}
__global__ void k_init_tiles_sfq(char* pool, uint64_t stride, uint64_t n, uint32_t ntiles) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (uint32_t t = 0; t < ntiles; ++t) gcq::sfq::init_slice((gcq::sfq::Queue*)(pool + t * stride), n, 0, gid, st);
}
__global__ void k_init_tiles_ymc(char* pool, uint64_t stride, char* hpool, uint64_t hstride, uint64_t cells,
                                 uint32_t nthreads, uint32_t ntiles) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (uint32_t t = 0; t < ntiles; ++t)
    gcq::ymc::init_slice((gcq::ymc::Queue*)(pool + t * stride), (gcq::ymc::Handle*)(hpool + t * hstride), cells,
                         nthreads, 16, gid, st);
}

// The camera origin's object-test terms (PrimaryTerms), one thread per
// object.
__global__ void k_primary_terms(const gcq_rt_sphere* sph, int32_t ns, const gcq_rt_plane* pl, int32_t np, V3 o,
                                PrimaryTerms* pre) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < ns) pre[i] = sphere_terms(o, sph[i]);
  else if (i < ns + np) pre[i] = plane_terms(o, pl[i - ns]);
}

}  // namespace rtk

namespace {

void fix_desc(gcq_rt_scene* s) {
  s->desc.spheres = s->sph.data();
  s->desc.nspheres = (int32_t)s->sph.size();
  s->desc.planes = s->pl.data();
  s->desc.nplanes = (int32_t)s->pl.size();
}

void free_ctx(gcq_rt_scene* s) {
  for (void* p : {(void*)s->qpool, (void*)s->hpool, s->adapters, (void*)s->recs, (void*)s->img, (void*)s->counters,
                  (void*)s->err})
    if (p) (void)hipFree(p);
  s->qpool = s->hpool = nullptr;
  s->adapters = nullptr;
  s->recs = nullptr;
  s->img = nullptr;
  s->counters = nullptr;
  s->err = nullptr;
  s->recs_cap = s->img_cap = 0;
  s->counters_cap = 0;
  s->kind = -1;
  s->ntiles = s->cap = s->threads = 0;
}

// Per-tile queues able to hold every pixel of a tile at once.
int ensure_queues(gcq_rt_scene* s, int kind, uint32_t ntiles, uint32_t tile_px, uint32_t threads) {
  const uint32_t cap = gcqf::pow2_at_least(std::max(tile_px, threads));
  if (kind == s->kind && ntiles <= s->ntiles && cap <= s->cap && threads <= s->threads && s->qpool) return GCQ_OK;
  for (void* p : {(void*)s->qpool, (void*)s->hpool, s->adapters})
    if (p) (void)hipFree(p);
  s->qpool = s->hpool = nullptr;
  s->adapters = nullptr;
  uint64_t qb = 0;
  switch (kind) {
    case GCQ_GWFQ:
      GCQ_REQUIRE(cap <= (1u << 25), "tile too large for G-WFQ");
      qb = gcq::gwf::ring_bytes(cap, threads);
      break;
    case GCQ_GLFQ: qb = gcq::lfq::ring_bytes(cap); break;
    case GCQ_SFQ: qb = gcq::sfq::queue_bytes(cap); break;
    case GCQ_WFQ:
      // Every bounce enqueue takes a cell and idle lanes may burn one per
      // attempt; the pool is sized for (bounces + 2) per pixel plus slack.
      s->ymc_cells = 8ull * tile_px + 256ull * threads + (1ull << 16);
      qb = gcq::ymc::queue_bytes(s->ymc_cells);
      break;
    default: GCQ_REQUIRE(false, "unknown queue kind %d", kind);
  }
  qb = (qb + 255) & ~255ull;
  GCQ_HIP(hipMalloc(&s->qpool, qb * ntiles));
  s->qstride = qb;
  if (kind == GCQ_WFQ) {
    s->hstride = sizeof(gcq::ymc::Handle) * (uint64_t)threads;
    GCQ_HIP(hipMalloc(&s->hpool, s->hstride * ntiles));
  }
  GCQ_HIP(hipMalloc(&s->adapters, 64ull * ntiles));
  // adapter array
  std::vector<char> a(64ull * ntiles, 0);
  for (uint32_t t = 0; t < ntiles; ++t) {
    char* qp = s->qpool + t * qb;
    switch (kind) {
      case GCQ_GWFQ: ((gcqf::FGwf*)a.data())[t] = gcqf::FGwf{(gcq::gwf::Ring*)qp}; break;
      case GCQ_GLFQ: ((gcqf::FLfq*)a.data())[t] = gcqf::FLfq{(gcq::lfq::Ring*)qp}; break;
      case GCQ_SFQ: ((gcqf::FSfq*)a.data())[t] = gcqf::FSfq{(gcq::sfq::Queue*)qp}; break;
      case GCQ_WFQ:
        ((gcqf::FYmc*)a.data())[t] = gcqf::FYmc{(gcq::ymc::Queue*)qp, (gcq::ymc::Handle*)(s->hpool + t * s->hstride)};
        break;
    }
  }
  GCQ_HIP(hipMemcpy(s->adapters, a.data(), a.size(), hipMemcpyHostToDevice));
  s->kind = kind;
  s->ntiles = ntiles;
  s->cap = cap;
  s->threads = threads;
  return GCQ_OK;
}

void reset_queues(gcq_rt_scene* s, uint32_t ntiles) {
  const uint32_t grid = gcqf::grid_of(2ull * s->cap);
  switch (s->kind) {
    case GCQ_GWFQ: {
      gcq::gwf::InitParams p{s->cap, s->threads, 0, 64, 128, 256, 0};
      hipLaunchKernelGGL(rtk::k_init_tiles_gwf, dim3(grid), dim3(256), 0, s->stream, s->qpool, s->qstride, p, ntiles);
      break;
    }
    case GCQ_GLFQ:
      hipLaunchKernelGGL(rtk::k_init_tiles_lfq, dim3(grid), dim3(256), 0, s->stream, s->qpool, s->qstride, s->cap,
                         ntiles);
      break;
    case GCQ_SFQ:
      hipLaunchKernelGGL(rtk::k_init_tiles_sfq, dim3(grid), dim3(256), 0, s->stream, s->qpool, s->qstride,
                         (uint64_t)s->cap, ntiles);
      break;
    case GCQ_WFQ:
      for (uint32_t t = 0; t < ntiles; ++t)
        hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(2 * s->ymc_cells)), dim3(256), 0, s->stream,
                           (uint64_t*)(s->qpool + t * s->qstride + gcq::ymc::kHdrBytes), 2 * s->ymc_cells);
      hipLaunchKernelGGL(rtk::k_init_tiles_ymc, dim3(gcqf::grid_of(s->threads)), dim3(256), 0, s->stream, s->qpool,
                         s->qstride, s->hpool, s->hstride, s->ymc_cells, s->threads, ntiles);
      break;
  }
}

// splitmix64 stream for the random scene.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void set3(float* d, float x, float y, float z) {
  d[0] = x;
  d[1] = y;
  d[2] = z;
}

}  // namespace

int rtk::scene_upload(gcq_rt_scene* s) {
  if (s->stream) return GCQ_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  GCQ_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  GCQ_HIP(hipMalloc(&s->d_sph, sizeof(gcq_rt_sphere) * std::max<size_t>(1, s->sph.size())));
  GCQ_HIP(hipMalloc(&s->d_pl, sizeof(gcq_rt_plane) * std::max<size_t>(1, s->pl.size())));
  if (!s->sph.empty())
    GCQ_HIP(hipMemcpy(s->d_sph, s->sph.data(), sizeof(gcq_rt_sphere) * s->sph.size(), hipMemcpyHostToDevice));
  if (!s->pl.empty())
    GCQ_HIP(hipMemcpy(s->d_pl, s->pl.data(), sizeof(gcq_rt_plane) * s->pl.size(), hipMemcpyHostToDevice));
  // The camera's primary terms; the renderers use them only once they are
  // complete (d_pre stays null otherwise, and every ray takes the full test).
  const int32_t nobj = (int32_t)(s->sph.size() + s->pl.size());
  if (nobj) {
    rtk::PrimaryTerms* pre = nullptr;
    GCQ_HIP(hipMalloc(&pre, sizeof(rtk::PrimaryTerms) * (size_t)nobj));
    const float* o = s->desc.cam_o;
    hipLaunchKernelGGL(rtk::k_primary_terms, dim3((nobj + 255) / 256), dim3(256), 0, s->stream, s->d_sph,
                       (int32_t)s->sph.size(), s->d_pl, (int32_t)s->pl.size(), rtk::v3(o[0], o[1], o[2]), pre);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
      (void)hipFree(pre);
      gcqh::set_error("primary-ray terms kernel failed");
      return GCQ_EHIP;
    }
    s->d_pre = pre;
  }
  return GCQ_OK;
}

rtk::Scene rtk::scene_view(const gcq_rt_scene* s) {
  rtk::Scene sc;
  sc.sph = s->d_sph;
  sc.pl = s->d_pl;
  sc.pre = s->d_pre;
  sc.ns = (int32_t)s->sph.size();
  sc.np = (int32_t)s->pl.size();
  const gcq_rt_scene_desc& d = s->desc;
  auto hn = [](const float* v) { return rtk::norm(rtk::v3(v[0], v[1], v[2])); };
  sc.cam_o = rtk::v3(d.cam_o[0], d.cam_o[1], d.cam_o[2]);
  sc.cam_f = hn(d.cam_f);  // rt_final.cpp:851-859 normalises on the host
  sc.cam_r = hn(d.cam_r);
  sc.cam_u = hn(d.cam_u);
  sc.light_raw = rtk::v3(d.light[0], d.light[1], d.light[2]);
  sc.ambient = rtk::v3(d.ambient[0], d.ambient[1], d.ambient[2]);
  return sc;
}

extern "C" void gcq_rt_config_default(gcq_rt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->kind = GCQ_GLFQ;
  c->threads_per_tile = 1024;  // rt_final.cpp:713
  c->block = 256;              // TPB, :320
  c->max_depth = 4;            // :712
  c->width = 1280;             // IMG_W/H, :313-314
  c->height = 720;
  c->tile_w = 160;             // :317-318
  c->tile_h = 90;
}

extern "C" int gcq_rt_scene_create(const gcq_rt_scene_desc* d, gcq_rt_scene** out) {
  GCQ_REQUIRE(d && out && d->nspheres >= 0 && d->nplanes >= 0, "bad scene");
  GCQ_REQUIRE((d->spheres || d->nspheres == 0) && (d->planes || d->nplanes == 0), "bad scene arrays");
  gcq_rt_scene* s = new gcq_rt_scene();
  s->sph.assign(d->spheres, d->spheres + d->nspheres);
  s->pl.assign(d->planes, d->planes + d->nplanes);
  s->desc = *d;
  fix_desc(s);
  *out = s;
  return GCQ_OK;
}

extern "C" int gcq_rt_scene_preset(int32_t preset, int32_t nspheres, uint64_t seed, gcq_rt_scene** out) {
  GCQ_REQUIRE(out, "null out");
  gcq_rt_scene* s = new gcq_rt_scene();
  gcq_rt_scene_desc& d = s->desc;
  std::memset(&d, 0, sizeof d);
  auto sphere = [&](float x, float y, float z, float r, float ar, float ag, float ab, float refl) {
    gcq_rt_sphere v;
    set3(v.c, x, y, z);
    v.r = r;
    set3(v.albedo, ar, ag, ab);
    v.refl = refl;
    s->sph.push_back(v);
  };
  auto plane = [&](float nx, float ny, float nz, float dd, float ar, float ag, float ab, float refl) {
    gcq_rt_plane v;
    set3(v.n, nx, ny, nz);
    v.d = dd;
    set3(v.albedo, ar, ag, ab);
    v.refl = refl;
    s->pl.push_back(v);
  };
  set3(d.light, -0.5f, 1.f, 0.2f);  // rt_final.cpp:560
  set3(d.ambient, 0.05f, 0.05f, 0.05f);
  set3(d.cam_r, 1.f, 0.f, 0.f);     // :856-857
  set3(d.cam_u, 0.f, 1.f, 0.f);
  if (preset == GCQ_RT_COMPLEX) {
    // build_scene_complex (rt_final.cpp:650-671): floor, 10 x 10 grid of
    // small spheres cycling through 7 materials, three large spheres.
    plane(0, 1, 0, 1.f, 0.75f, 0.75f, 0.75f, 0.05f);
    const int gx = 10, gz = 10;
    const float sp = 0.55f, rad = 0.22f, base_y = 0.25f;
    static const float kRefl[7] = {0.9f, 0.6f, 0.3f, 0.1f, 0.f, 0.f, 0.f};
    static const float kAlb[7][3] = {{.9f, .9f, .9f}, {.9f, .5f, .2f}, {.2f, .8f, .9f}, {.7f, .2f, .8f},
                                     {.2f, .9f, .3f}, {.9f, .2f, .2f}, {.8f, .8f, .3f}};
    for (int iz = 0; iz < gz; ++iz)
      for (int ix = 0; ix < gx; ++ix) {
        const float x = (ix - (gx - 1) * 0.5f) * sp;
        const float z = -2.f - iz * sp;
        const int m = (ix + iz * 7) % 7;
        sphere(x, base_y, z, rad, kAlb[m][0], kAlb[m][1], kAlb[m][2], kRefl[m]);
      }
    sphere(-1.2f, 0.6f, -3.5f, 0.6f, .9f, .3f, .3f, 0.7f);
    sphere(1.4f, 0.9f, -4.2f, 0.9f, .3f, .9f, .4f, 0.2f);
    sphere(0.0f, 1.3f, -5.5f, 1.3f, .4f, .6f, .95f, 0.0f);
    set3(d.cam_o, 0.f, 0.3f, 0.5f);  // :850-852
    set3(d.cam_f, 0.f, -0.2f, -1.f);
  } else if (preset == GCQ_RT_CORNELL) {
    // build_scene_cornell (rt_final.cpp:673-684).
    sphere(-0.6f, 0.5f, -2.5f, 0.5f, .9f, .3f, .3f, 0.85f);
    sphere(0.6f, 0.6f, -2.8f, 0.6f, .95f, .95f, .95f, 0.95f);
    plane(0, 1, 0, 1.f, .9f, .9f, .9f, 0.3f);
    plane(0, 0, 1, 5.f, .6f, .65f, .8f, 0.7f);
    plane(1, 0, 0, 2.f, .8f, .2f, .2f, 0.6f);
    plane(-1, 0, 0, 2.f, .2f, .8f, .3f, 0.6f);
    plane(0, -1, 0, 2.5f, .95f, .95f, .95f, 0.2f);
    set3(d.cam_o, 0.f, 0.5f, 1.5f);  // :853-855
    set3(d.cam_f, 0.f, -0.1f, -1.f);
  } else if (preset == GCQ_RT_RANDOM) {
    // n spheres resting on the ground plane y = -1: radius in [0.2, 1.5],
    // 70% Lambertian (refl 0), 30% mirror (refl 0.8), albedo in [0.15, 0.95].
    if (nspheres < 0 || nspheres > 1 << 20) {
      delete s;
      GCQ_REQUIRE(false, "bad sphere count %d", nspheres);
    }
    Rng g{seed};
    plane(0, 1, 0, 1.f, 0.8f, 0.8f, 0.8f, 0.f);
    for (int32_t i = 0; i < nspheres; ++i) {
      const float r = (float)(0.2 + 1.3 * g.u());
      const float x = (float)(-5.0 + 10.0 * g.u());
      const float z = (float)(-1.5 - 10.5 * g.u());
      const float ar = (float)(0.15 + 0.8 * g.u()), ag = (float)(0.15 + 0.8 * g.u()), ab = (float)(0.15 + 0.8 * g.u());
      const float refl = g.u() < 0.3 ? 0.8f : 0.f;
      sphere(x, -1.f + r, z, r, ar, ag, ab, refl);
    }
    set3(d.cam_o, 0.f, 0.8f, 2.0f);
    set3(d.cam_f, 0.f, -0.2f, -1.f);
  } else {
    delete s;
    GCQ_REQUIRE(false, "unknown scene preset %d", preset);
  }
  fix_desc(s);
  *out = s;
  return GCQ_OK;
}

extern "C" int gcq_rt_scene_get(const gcq_rt_scene* s, gcq_rt_scene_desc* d) {
  GCQ_REQUIRE(s && d, "null argument");
  *d = s->desc;
  return GCQ_OK;
}

extern "C" int gcq_rt_scene_destroy(gcq_rt_scene* s) {
  if (!s) return GCQ_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_ctx(s);
  rtk::free_compact(s);
  if (s->d_sph) (void)hipFree(s->d_sph);
  if (s->d_pre) (void)hipFree(s->d_pre);
  if (s->d_pl) (void)hipFree(s->d_pl);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GCQ_OK;
}

extern "C" int gcq_rt_render(gcq_rt_scene* s, const gcq_rt_config* cfg, float* img_out, gcq_rt_result* res) {
  GCQ_REQUIRE(s && cfg && res, "null argument");
  GCQ_REQUIRE(cfg->kind >= 0 && cfg->kind <= 3, "unknown queue kind %d", cfg->kind);
  GCQ_REQUIRE(cfg->block >= 64 && cfg->block <= 1024 && cfg->block % 64 == 0, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->threads_per_tile >= cfg->block && cfg->threads_per_tile % cfg->block == 0,
              "threads_per_tile must be a positive multiple of block");
  GCQ_REQUIRE(cfg->width > 0 && cfg->height > 0 && cfg->tile_w > 0 && cfg->tile_h > 0 && cfg->max_depth >= 0,
              "bad image geometry");
  const int32_t w = cfg->w ? cfg->w : cfg->width, h = cfg->h ? cfg->h : cfg->height;
  GCQ_REQUIRE(cfg->x0 >= 0 && cfg->y0 >= 0 && w > 0 && h > 0 && cfg->x0 + w <= cfg->width && cfg->y0 + h <= cfg->height,
              "region outside the image");
  GCQ_REQUIRE((int64_t)cfg->tile_w * cfg->tile_h < (1ll << 30), "tile too large");
  std::memset(res, 0, sizeof *res);
  int rc = rtk::scene_upload(s);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const int32_t tiles_x = (w + cfg->tile_w - 1) / cfg->tile_w, tiles_y = (h + cfg->tile_h - 1) / cfg->tile_h;
  const uint32_t ntiles = (uint32_t)(tiles_x * tiles_y);
  GCQ_REQUIRE(ntiles <= 65535, "too many tiles (%u)", ntiles);
  const uint32_t tile_px = (uint32_t)(std::min(cfg->tile_w, w) * std::min(cfg->tile_h, h));
  rc = ensure_queues(s, cfg->kind, ntiles, tile_px, cfg->threads_per_tile);
  if (rc) return rc;
  const uint64_t npx = (uint64_t)w * h;
  if (npx > s->recs_cap) {
    if (s->recs) (void)hipFree(s->recs);
    if (s->img) (void)hipFree(s->img);
    s->recs = nullptr;
    s->img = nullptr;
    s->recs_cap = 0;
    GCQ_HIP(hipMalloc(&s->recs, sizeof(rtk::RayRec) * npx));
    GCQ_HIP(hipMalloc(&s->img, 12ull * npx));
    s->recs_cap = npx;
  }
  if (ntiles > s->counters_cap) {
    if (s->counters) (void)hipFree(s->counters);
    s->counters = nullptr;
    GCQ_HIP(hipMalloc(&s->counters, 2ull * rtk::kCounterStride * 8 * ntiles));
    s->counters_cap = ntiles;
  }
  if (!s->err) GCQ_HIP(hipMalloc(&s->err, 8));

  const rtk::Scene sc = rtk::scene_view(s);
  rtk::Job jb;
  jb.W = cfg->width;
  jb.H = cfg->height;
  jb.x0 = cfg->x0;
  jb.y0 = cfg->y0;
  jb.w = w;
  jb.h = h;
  jb.tile_w = cfg->tile_w;
  jb.tile_h = cfg->tile_h;
  jb.tiles_x = tiles_x;
  jb.max_depth = cfg->max_depth;
  jb.threads_per_tile = cfg->threads_per_tile;
  jb.recs = s->recs;
  jb.img = s->img;
  jb.alive = s->counters;
  jb.traced = s->counters + (uint64_t)rtk::kCounterStride * ntiles;
  jb.err = s->err;

  reset_queues(s, ntiles);
  hipLaunchKernelGGL(rtk::k_init_alive, dim3((ntiles + 255) / 256), dim3(256), 0, s->stream, jb.alive, jb.traced, jb,
                     ntiles);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(64), 0, s->stream, (uint64_t*)s->err, (uint64_t)1);
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  GCQ_HIP(hipEventRecord(e0, s->stream));
  const dim3 grid(cfg->threads_per_tile / cfg->block, ntiles);
  switch (s->kind) {
    case GCQ_GWFQ:
      hipLaunchKernelGGL(rtk::k_render<gcqf::FGwf>, grid, dim3(cfg->block), 0, s->stream,
                         (const gcqf::FGwf*)s->adapters, sc, jb);
      break;
    case GCQ_GLFQ:
      hipLaunchKernelGGL(rtk::k_render<gcqf::FLfq>, grid, dim3(cfg->block), 0, s->stream,
                         (const gcqf::FLfq*)s->adapters, sc, jb);
      break;
    case GCQ_SFQ:
      hipLaunchKernelGGL(rtk::k_render<gcqf::FSfq>, grid, dim3(cfg->block), 0, s->stream,
                         (const gcqf::FSfq*)s->adapters, sc, jb);
      break;
    case GCQ_WFQ:
      hipLaunchKernelGGL(rtk::k_render<gcqf::FYmc>, grid, dim3(cfg->block), 0, s->stream,
                         (const gcqf::FYmc*)s->adapters, sc, jb);
      break;
  }
  GCQ_HIP(hipGetLastError());
  GCQ_HIP(hipEventRecord(e1, s->stream));
  std::vector<uint64_t> cnt(2ull * rtk::kCounterStride * ntiles);
  uint32_t err = 0;
  GCQ_HIP(hipMemcpyAsync(cnt.data(), s->counters, cnt.size() * 8, hipMemcpyDeviceToHost, s->stream));
  GCQ_HIP(hipMemcpyAsync(&err, s->err, 4, hipMemcpyDeviceToHost, s->stream));
  if (img_out) GCQ_HIP(hipMemcpyAsync(img_out, s->img, 12ull * npx, hipMemcpyDeviceToHost, s->stream));
  GCQ_HIP(hipStreamSynchronize(s->stream));
  res->kernel_ms = gcqh::ms_between(e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->primary_rays = npx;
  for (uint32_t t = 0; t < ntiles; ++t) res->total_rays += cnt[(uint64_t)rtk::kCounterStride * (ntiles + t)];
  res->tiles = (int32_t)ntiles;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (err & 2u) {
    gcqh::set_error("ray tracing aborted: a tile's workers idled past the spin limit");
    return GCQ_ETIMEOUT;
  }
  if (err & 1u) {
    gcqh::set_error("ray tracing: a bounce ray could not be enqueued");
    return GCQ_EFULL;
  }
  return GCQ_OK;
}

// ======================================================================== image and CSV outputs
// 8-bit binary PPM with the reference's quantisation: each channel becomes
// (unsigned char)min(255, max(0, x * 255)), truncating (save_ppm,
// rt_final.cpp:689-700; the compaction baseline writes the same format).
extern "C" int gcq_rt_write_ppm(const float* img, int32_t w, int32_t h, const char* path) {
  GCQ_REQUIRE(img && path && w > 0 && h > 0, "bad image or path");
  FILE* f = fopen(path, "wb");
  GCQ_REQUIRE(f, "cannot open %s", path);
  fprintf(f, "P6\n%d %d\n255\n", w, h);
  std::vector<unsigned char> row(3ull * w);
  for (int32_t y = 0; y < h; ++y) {
    const float* src = img + 3ll * w * y;
    for (int64_t i = 0; i < 3ll * w; ++i) row[i] = (unsigned char)fminf(255.f, fmaxf(0.f, src[i] * 255.f));
    if (fwrite(row.data(), 1, row.size(), f) != row.size()) {
      fclose(f);
      GCQ_REQUIRE(false, "short write to %s", path);
    }
  }
  fclose(f);
  return GCQ_OK;
}

// One result row appended to a CSV file, with its header when the file is new.
//  GCQ_RT_CSV_QUEUE: the queue renderer's rt_queue_perf.csv header
//    (rt_final.cpp:978-980), every column filled.  The reference's row omits
//    PRIMARY, GEN_MS and TRACE_MS (:992-999); here GEN_MS is 0 because ray
//    generation runs inside the one trace launch, and MRAYS is rays over
//    trace time (:938).
//  GCQ_RT_CSV_COMPACT: the compaction baseline's rt_performance.csv
//    (baselines/rt_compaction.cpp:849-864), THREADS = block size, MRAYS over
//    the frame time.
extern "C" int gcq_rt_append_csv(const char* path, int32_t style, const gcq_rt_csv_row* r) {
  GCQ_REQUIRE(path && r && r->res && r->gpu && r->queue, "null argument");
  GCQ_REQUIRE(style == GCQ_RT_CSV_QUEUE || style == GCQ_RT_CSV_COMPACT, "unknown CSV style %d", style);
  FILE* probe = fopen(path, "rb");
  const bool exists = probe != nullptr;
  if (probe) fclose(probe);
  FILE* f = fopen(path, "ab");
  GCQ_REQUIRE(f, "cannot open %s", path);
  const gcq_rt_result& x = *r->res;
  if (style == GCQ_RT_CSV_QUEUE) {
    if (!exists) fprintf(f, "GPU,QUEUE,SCENE,THREADS,BOUNCES,PRIMARY,TOTAL,GEN_MS,TRACE_MS,TOTAL_MS,MRAYS\n");
    const double mrays = x.kernel_ms > 0 ? (double)x.total_rays / 1e6 / (x.kernel_ms / 1e3) : 0.0;
    fprintf(f, "\"%s\",%s,%s,%u,%d,%llu,%llu,%g,%g,%g,%g\n", r->gpu, r->queue, r->scene ? r->scene : "", r->threads,
            r->bounces, (unsigned long long)x.primary_rays, (unsigned long long)x.total_rays, 0.0, x.kernel_ms,
            x.total_ms, mrays);
  } else {
    if (!exists) fprintf(f, "DEVICE,QUEUE,THREADS,TOTAL_RAYS,TOTAL_TIME_MS,TOTAL_MRAYS\n");
    const double mrays = x.kernel_ms > 0 ? (double)x.total_rays / 1e6 / (x.kernel_ms / 1e3) : 0.0;
    fprintf(f, "%s,%s,%u,%llu,%g,%g\n", r->gpu, r->queue, r->threads, (unsigned long long)x.total_rays, x.kernel_ms,
            mrays);
  }
  fclose(f);
  return GCQ_OK;
}
