// rt_shade.hpp — per-ray arithmetic of the ray tracer, shared by the render
// kernel (rt.hip) and the serving kernel (serve.hip).
//
// Follows rt_final.cpp: float3 helpers (:329-341), intersect_sphere /
// intersect_plane / trace_scene (:368-419), the camera ray of k_generate
// (:482-495) and one k_trace bounce (:572-628).  Including files turn
// floating-point contraction off so every operation is the plain IEEE float
// operation of the source expression, in source order.
#pragma once
#pragma clang fp contract(off)
#include "frontier.hpp"

namespace rtk {

struct V3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ V3 norm(V3 v) {
  const float d = sqrtf(dot(v, v));
  return d > 0.f ? v * (1.f / d) : v;
}
__host__ __device__ __forceinline__ V3 reflect(V3 v, V3 n) { return v - n * (2.f * dot(v, n)); }
__host__ __device__ __forceinline__ V3 clamp01(V3 c) {
  return v3(fminf(fmaxf(c.x, 0.f), 1.f), fminf(fmaxf(c.y, 0.f), 1.f), fminf(fmaxf(c.z, 0.f), 1.f));
}
__host__ __device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// A live bounce ray; also holds its pixel's running sum.
struct alignas(16) RayRec {
  float o[3];
  int32_t depth;
  float d[3];
  float pad0;
  float at[3];
  float pad1;
  float acc[3];
  float pad2;
};
static_assert(sizeof(RayRec) == 64, "ray record");

struct Ray {
  V3 o, d, at, acc;
  int32_t depth;
};

__device__ __forceinline__ void store_ray(RayRec* r, const Ray& x) {
  uint64_t* w = (uint64_t*)r;
  auto pk = [](float a, float b) {
    return (uint64_t)__float_as_uint(a) | ((uint64_t)__float_as_uint(b) << 32);
  };
  gcq::st_rlx(&w[0], pk(x.o.x, x.o.y));
  gcq::st_rlx(&w[1], pk(x.o.z, __int_as_float(x.depth)));
  gcq::st_rlx(&w[2], pk(x.d.x, x.d.y));
  gcq::st_rlx(&w[3], pk(x.d.z, 0.f));
  gcq::st_rlx(&w[4], pk(x.at.x, x.at.y));
  gcq::st_rlx(&w[5], pk(x.at.z, 0.f));
  gcq::st_rlx(&w[6], pk(x.acc.x, x.acc.y));
  gcq::st_rlx(&w[7], pk(x.acc.z, 0.f));
}
__device__ __forceinline__ Ray load_ray(const RayRec* r) {
  const uint64_t* w = (const uint64_t*)r;
  uint64_t a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = gcq::ld_rlx(&w[i]);
  auto lo = [](uint64_t x) { return __uint_as_float((uint32_t)x); };
  auto hi = [](uint64_t x) { return __uint_as_float((uint32_t)(x >> 32)); };
  Ray x;
  x.o = v3(lo(a[0]), hi(a[0]), lo(a[1]));
  x.depth = __float_as_int(hi(a[1]));
  x.d = v3(lo(a[2]), hi(a[2]), lo(a[3]));
  x.at = v3(lo(a[4]), hi(a[4]), lo(a[5]));
  x.acc = v3(lo(a[6]), hi(a[6]), lo(a[7]));
  return x;
}

// Terms of the object tests that depend only on the ray origin, computed
// once per scene for the camera origin that every primary ray starts from:
// per sphere oc = o - c and dot(oc, oc) - r * r, per plane dot(n, o) + d
// (in ocx).  They are the same float operations the tests perform, so a
// primary ray's hits are bit-identical with or without them.
struct PrimaryTerms {
  float ocx, ocy, ocz, c;
};

struct Scene {
  const gcq_rt_sphere* sph;
  const gcq_rt_plane* pl;
  const PrimaryTerms* pre;  // ns sphere terms, then np plane terms; null: none
  int32_t ns, np;
  V3 cam_o, cam_f, cam_r, cam_u;  // f, r, u already normalised on the host
  V3 light_raw, ambient;
};

// Nearest hit (trace_scene, rt_final.cpp:403-419): spheres in order, then
// planes; a later object replaces the best only if strictly nearer.
struct Hit {
  float t;
  int32_t kind, idx;  // kind 0 none, 1 sphere, 2 plane
};
// The scene's spheres and planes are uploaded once and no kernel writes
// them, so the tests read them through the constant address space: every
// lane tests the same object, and the compiler turns those uniform reads
// into scalar loads issued ahead of the arithmetic, where a global pointer
// gave one vector load and one wait per object.
typedef const __attribute__((address_space(4))) gcq_rt_sphere* ConstSphere;
typedef const __attribute__((address_space(4))) gcq_rt_plane* ConstPlane;
// One sphere test (intersect_sphere, rt_final.cpp:368-385) against the
// best hit so far.
__device__ __forceinline__ void sphere_test(Hit& h, int32_t i, V3 o, V3 d, float cx, float cy, float cz, float r) {
  const V3 oc = o - v3(cx, cy, cz);
  const float b = dot(oc, d);
  const float c = dot(oc, oc) - r * r;
  const float disc = b * b - c;
  if (disc >= 0.f) {
    const float t = -b - sqrtf(disc);
    if (t > 1e-4f && t < h.t) h = Hit{t, 1, i};
  }
}
__device__ __forceinline__ Hit nearest(const Scene& sc, V3 o, V3 d) {
  Hit h{1e30f, 0, -1};
  const ConstSphere sph = (ConstSphere)sc.sph;
  const ConstPlane pl = (ConstPlane)sc.pl;
  // Spheres in groups of 8: the group's centres and radii are loaded first,
  // so their scalar loads are in flight together rather than one wait per
  // sphere; the tests then run in sphere order as before.
  constexpr int32_t kGroup = 8;
  int32_t i = 0;
  for (; i + kGroup <= sc.ns; i += kGroup) {
    float cx[kGroup], cy[kGroup], cz[kGroup], rr[kGroup];
#pragma unroll
    for (int32_t k = 0; k < kGroup; ++k) {
      cx[k] = sph[i + k].c[0];
      cy[k] = sph[i + k].c[1];
      cz[k] = sph[i + k].c[2];
      rr[k] = sph[i + k].r;
    }
#pragma unroll
    for (int32_t k = 0; k < kGroup; ++k) sphere_test(h, i + k, o, d, cx[k], cy[k], cz[k], rr[k]);
  }
  for (; i < sc.ns; ++i) sphere_test(h, i, o, d, sph[i].c[0], sph[i].c[1], sph[i].c[2], sph[i].r);
  for (int32_t j = 0; j < sc.np; ++j) {
    const auto& p = pl[j];
    const V3 n = v3(p.n[0], p.n[1], p.n[2]);
    const float denom = dot(n, d);
    if (fabsf(denom) > 1e-5f) {
      const float t = -(dot(n, o) + p.d) / denom;
      if (t > 1e-4f && t < h.t) h = Hit{t, 2, j};
    }
  }
  return h;
}

__device__ __forceinline__ PrimaryTerms sphere_terms(V3 o, const gcq_rt_sphere& s) {
  const V3 oc = o - v3(s.c[0], s.c[1], s.c[2]);
  return PrimaryTerms{oc.x, oc.y, oc.z, dot(oc, oc) - s.r * s.r};
}
__device__ __forceinline__ PrimaryTerms plane_terms(V3 o, const gcq_rt_plane& p) {
  return PrimaryTerms{dot(v3(p.n[0], p.n[1], p.n[2]), o) + p.d, 0.f, 0.f, 0.f};
}

// nearest() for a ray from the camera origin, from the scene's primary
// terms: per sphere b = dot(oc, d) and disc = b * b - c remain (7 of the
// 17 operations), per plane the denominator.
typedef const __attribute__((address_space(4))) PrimaryTerms* ConstTerms;
__device__ __forceinline__ void sphere_test_pre(Hit& h, int32_t i, V3 d, PrimaryTerms q) {
  const float b = dot(v3(q.ocx, q.ocy, q.ocz), d);
  const float disc = b * b - q.c;
  if (disc >= 0.f) {
    const float t = -b - sqrtf(disc);
    if (t > 1e-4f && t < h.t) h = Hit{t, 1, i};
  }
}
__device__ __forceinline__ Hit nearest_primary(const Scene& sc, V3 d) {
  Hit h{1e30f, 0, -1};
  const ConstTerms pre = (ConstTerms)sc.pre;
  const ConstPlane pl = (ConstPlane)sc.pl;
  constexpr int32_t kGroup = 8;
  int32_t i = 0;
  for (; i + kGroup <= sc.ns; i += kGroup) {
    PrimaryTerms q[kGroup];
#pragma unroll
    for (int32_t k = 0; k < kGroup; ++k) q[k] = PrimaryTerms{pre[i + k].ocx, pre[i + k].ocy, pre[i + k].ocz, pre[i + k].c};
#pragma unroll
    for (int32_t k = 0; k < kGroup; ++k) sphere_test_pre(h, i + k, d, q[k]);
  }
  for (; i < sc.ns; ++i) sphere_test_pre(h, i, d, PrimaryTerms{pre[i].ocx, pre[i].ocy, pre[i].ocz, pre[i].c});
  for (int32_t j = 0; j < sc.np; ++j) {
    const auto& p = pl[j];
    const float denom = dot(v3(p.n[0], p.n[1], p.n[2]), d);
    if (fabsf(denom) > 1e-5f) {
      const float t = -pre[sc.ns + j].ocx / denom;
      if (t > 1e-4f && t < h.t) h = Hit{t, 2, j};
    }
  }
  return h;
}

// Primary ray of pixel (px, py) (k_generate, rt_final.cpp:482-495).
__device__ __forceinline__ Ray primary(const Scene& sc, int32_t W, int32_t H, int32_t px, int32_t py) {
  const float aspect = (float)W / (float)H;
  const float u = ((px + 0.5f) / W - 0.5f) * 2.f * aspect;
  const float v = -((py + 0.5f) / H - 0.5f) * 2.f;
  Ray r;
  r.o = sc.cam_o;
  r.d = norm(sc.cam_f + sc.cam_r * u + sc.cam_u * v);
  r.at = v3(1.f, 1.f, 1.f);
  r.acc = v3(0.f, 0.f, 0.f);
  r.depth = 0;
  return r;
}

// One bounce of ray r (k_trace body, rt_final.cpp:572-628).  Returns true
// and updates r in place if the chain continues.
__device__ __forceinline__ bool shade(const Scene& sc, int32_t max_depth, V3 light, Ray& r) {
  // depth 0: a primary ray, whose origin is the camera's (primary())
  const Hit h = sc.pre && r.depth == 0 ? nearest_primary(sc, r.d) : nearest(sc, r.o, r.d);
  if (h.kind == 0) {
    const float t = 0.5f * (r.d.y + 1.f);
    const V3 sky = v3(1.f, 1.f, 1.f) * (1.f - t) + v3(0.5f, 0.7f, 1.f) * t;
    r.acc = r.acc + clamp01(sky * r.at);
    return false;
  }
  const V3 p = r.o + r.d * h.t;
  V3 n, albedo;
  float refl;
  if (h.kind == 1) {
    const gcq_rt_sphere& s = sc.sph[h.idx];
    n = norm(p - ld3(s.c));
    albedo = ld3(s.albedo);
    refl = s.refl;
  } else {
    const gcq_rt_plane& q = sc.pl[h.idx];
    n = ld3(q.n);
    albedo = ld3(q.albedo);
    refl = q.refl;
  }
  const float ndl = fmaxf(0.f, dot(n, light));
  const V3 direct = sc.ambient + albedo * ndl;
  const V3 contrib = direct * r.at;
  if (refl > 0.f && r.depth < max_depth) {
    const V3 nr = clamp01(contrib * (1.f - refl));
    if (nr.x > 0.f || nr.y > 0.f || nr.z > 0.f) r.acc = r.acc + nr;
    r.o = p + n * 1e-4f;
    r.d = norm(reflect(r.d, n));
    r.depth = r.depth + 1;
    r.at = r.at * albedo * refl;
    return true;
  }
  r.acc = r.acc + clamp01(contrib);
  return false;
}

}  // namespace rtk
