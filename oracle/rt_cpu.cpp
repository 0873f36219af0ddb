// rt_cpu.cpp — ORACLE (test infrastructure only).  Scalar CPU restatement of
// the reference ray tracer's image: for each pixel, the camera ray of
// k_generate (rt_final.cpp:482-495), then per bounce the k_trace body
// (:572-628) with trace_scene / intersect_sphere / intersect_plane
// (:368-419) and the float3 helpers (:329-341); contributions are summed
// per pixel in bounce order, which is what accum_pixel's atomics (:424-429)
// produce because each pixel has exactly one live ray at a time.  Built
// with -ffp-contract=off: every operation is the plain IEEE float
// operation of the source expression, in source order.
// -DPLANT_RT_ATTEN_NO_ALBEDO builds a copy with a planted mistake in the
// reflection attenuation, which the closed-form tests must reject.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

struct F3 {
  float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 scale(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 mulv(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 normalize3(F3 v) {  // rt_final.cpp:335-338
  float d = std::sqrt(dot3(v, v));
  return d > 0.f ? scale(v, 1.f / d) : v;
}
inline F3 reflect3(F3 v, F3 n) { return sub(v, scale(n, 2.f * dot3(v, n))); }
inline float c01(float x) { return std::fmin(std::fmax(x, 0.f), 1.f); }
inline F3 clamp3(F3 c) { return f3(c01(c.x), c01(c.y), c01(c.z)); }
inline F3 at3(const float* p) { return f3(p[0], p[1], p[2]); }

struct Obj {  // sphere {c, r, albedo, refl} or plane {n, d, albedo, refl}: 8 floats
  const float* f;
};

struct RHit {
  bool hit;
  float t;
  F3 p, n, albedo;
  float refl;
};

RHit trace(const float* sph, int ns, const float* pl, int np, F3 o, F3 d) {
  RHit best{false, 1e30f, {}, {}, {}, 0.f};
  for (int i = 0; i < ns; ++i) {  // intersect_sphere, rt_final.cpp:368-389
    const float* s = sph + 8 * i;
    F3 oc = sub(o, at3(s));
    float b = dot3(oc, d);
    float c = dot3(oc, oc) - s[3] * s[3];
    float disc = b * b - c;
    if (disc >= 0.f) {
      float t = -b - std::sqrt(disc);
      if (t > 1e-4f && t < best.t) {
        best.hit = true;
        best.t = t;
        best.p = add(o, scale(d, t));
        best.n = normalize3(sub(best.p, at3(s)));
        best.albedo = at3(s + 4);
        best.refl = s[7];
      }
    }
  }
  for (int i = 0; i < np; ++i) {  // intersect_plane, :391-408
    const float* q = pl + 8 * i;
    F3 n = at3(q);
    float denom = dot3(n, d);
    if (std::fabs(denom) > 1e-5f) {
      float t = -(dot3(n, o) + q[3]) / denom;
      if (t > 1e-4f && t < best.t) {
        best.hit = true;
        best.t = t;
        best.p = add(o, scale(d, t));
        best.n = n;
        best.albedo = at3(q + 4);
        best.refl = q[7];
      }
    }
  }
  return best;
}

}  // namespace

extern "C" int orc_rt_render(const float* sph, int32_t ns, const float* pl, int32_t np, const float* cam12,
                             const float* light3, const float* ambient3, int32_t W, int32_t H, int32_t x0, int32_t y0,
                             int32_t w, int32_t h, int32_t max_depth, int32_t threads, float* img,
                             uint64_t* rays_traced) {
  if (W <= 0 || H <= 0 || w <= 0 || h <= 0 || x0 < 0 || y0 < 0 || x0 + w > W || y0 + h > H) return -1;
  const F3 cam_o = at3(cam12);
  const F3 cam_f = normalize3(at3(cam12 + 3));  // rt_final.cpp:851-859
  const F3 cam_r = normalize3(at3(cam12 + 6));
  const F3 cam_u = normalize3(at3(cam12 + 9));
  const F3 light = normalize3(at3(light3));     // :560
  const F3 ambient = at3(ambient3);             // :561
  const float aspect = (float)W / (float)H;     // :480
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::vector<uint64_t> counts(threads, 0);
  auto rows = [&](int t) {
    uint64_t traced = 0;
    for (int32_t ly = t; ly < h; ly += threads) {
      for (int32_t lx = 0; lx < w; ++lx) {
        const int32_t px = x0 + lx, py = y0 + ly;
        const float u = ((px + 0.5f) / W - 0.5f) * 2.f * aspect;  // :487-488
        const float v = -((py + 0.5f) / H - 0.5f) * 2.f;
        F3 o = cam_o;
        F3 d = normalize3(add(add(cam_f, scale(cam_r, u)), scale(cam_u, v)));
        F3 atten = f3(1.f, 1.f, 1.f);
        F3 pix = f3(0.f, 0.f, 0.f);
        for (int32_t depth = 0;; ++depth) {
          ++traced;
          const RHit hit = trace(sph, ns, pl, np, o, d);
          if (!hit.hit) {  // :622-627
            const float t = 0.5f * (d.y + 1.f);
            const F3 sky = add(scale(f3(1.f, 1.f, 1.f), 1.f - t), scale(f3(0.5f, 0.7f, 1.f), t));
            pix = add(pix, clamp3(mulv(sky, atten)));
            break;
          }
          const float ndotl = std::fmax(0.f, dot3(hit.n, light));  // :577-579
          const F3 direct = add(ambient, scale(hit.albedo, ndotl));
          const F3 contrib = mulv(direct, atten);
          if (hit.refl > 0.f && depth < max_depth) {  // :581-609
            const F3 nr = clamp3(scale(contrib, 1.f - hit.refl));
            if (nr.x > 0.f || nr.y > 0.f || nr.z > 0.f) pix = add(pix, nr);
            o = add(hit.p, scale(hit.n, 1e-4f));
            d = normalize3(reflect3(d, hit.n));
#ifdef PLANT_RT_ATTEN_NO_ALBEDO
            atten = scale(atten, hit.refl);  // planted mistake: the albedo factor dropped
#else
            atten = scale(mulv(atten, hit.albedo), hit.refl);
#endif
            continue;
          }
          pix = add(pix, clamp3(contrib));  // :610-614
          break;
        }
        float* out = img + 3ll * ((int64_t)ly * w + lx);
        out[0] = pix.x;
        out[1] = pix.y;
        out[2] = pix.z;
      }
    }
    counts[t] = traced;
  };
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) ts.emplace_back(rows, t);
  for (auto& x : ts) x.join();
  uint64_t total = 0;
  for (uint64_t c : counts) total += c;
  if (rays_traced) *rays_traced = total;
  return 0;
}
