// graph.cpp — synthetic CSR graphs for the BFS workloads (host side).
//
// CSR int32 (offsets[V+1], indices[E]), undirected (every edge stored in both
// directions), neighbours sorted ascending, no duplicates, no self loops —
// the form bfs_bench.cpp:181-200 builds (csr_from_edges after sort+unique).
//
// lattice(R, C, keep, diag, seed): the R x C 4-neighbour grid of
//   bfs_bench.cpp:245-256 where each lattice edge is kept with probability
//   `keep` and each vertex (r,c) adds, with probability `diag`, one short
//   diagonal to (r+1, c+1) or (r+1, c-1).  All draws are a splitmix64 hash of
//   (seed, edge or vertex id), so both directions agree and the CSR is built
//   directly in sorted order without sorting.
// rmat(scale, edge_factor, a, b, c, seed): 2^scale vertices, edge_factor *
//   2^scale generated edges, each endpoint pair drawn by recursive quadrant
//   choice with probabilities (a, b, c, 1-a-b-c); symmetrised, self loops and
//   duplicates removed (radix sort of (u,v) keys).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "graph.hpp"

namespace gcqg {

static inline uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static inline uint64_t draw(uint64_t seed, uint64_t salt, uint64_t key) {
  return mix64(mix64(seed ^ (salt * 0xD6E8FEB86659FD93ull)) ^ key);
}
static inline bool bern(uint64_t h, double p) {
  return (double)(h >> 11) * (1.0 / 9007199254740992.0) < p;
}

static int worker_count() {
  unsigned n = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(n, 16u));
}

template <class F>
static void parallel_rows(int64_t n, F f) {
  const int nt = worker_count();
  std::vector<std::thread> ts;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=]() { f(b, e); });
  }
  for (auto& x : ts) x.join();
}

bool lattice(int32_t R, int32_t C, double keep, double diag, uint64_t seed, Csr* g) {
  if (R <= 0 || C <= 0 || (int64_t)R * C >= (1ll << 31)) return false;
  const int64_t V = (int64_t)R * C;
  auto hkeep = [&](int64_t u, int dir) {  // dir 0: (r,c)-(r,c+1), 1: (r,c)-(r+1,c)
    return bern(draw(seed, 1, (uint64_t)(2 * u + dir)), keep);
  };
  auto hdiag = [&](int64_t u, int* dir) {  // diagonal from u to (r+1, c+dir)
    const uint64_t h = draw(seed, 2, (uint64_t)u);
    *dir = (h & 1) ? 1 : -1;
    return bern(h, diag);
  };
  // Neighbour enumeration in ascending id order.
  auto visit = [&](int64_t u, auto&& emit) {
    const int32_t r = (int32_t)(u / C), c = (int32_t)(u % C);
    int dir = 0;
    if (r > 0 && c > 0 && hdiag(u - C - 1, &dir) && dir == 1) emit(u - C - 1);
    if (r > 0 && hkeep(u - C, 1)) emit(u - C);
    if (r > 0 && c + 1 < C && hdiag(u - C + 1, &dir) && dir == -1) emit(u - C + 1);
    if (c > 0 && hkeep(u - 1, 0)) emit(u - 1);
    if (c + 1 < C && hkeep(u, 0)) emit(u + 1);
    const bool d = r + 1 < R && hdiag(u, &dir);
    if (d && dir == -1 && c > 0) emit(u + C - 1);
    if (r + 1 < R && hkeep(u, 1)) emit(u + C);
    if (d && dir == 1 && c + 1 < C) emit(u + C + 1);
  };
  g->nv = (int32_t)V;
  g->offsets.assign(V + 1, 0);
  parallel_rows(V, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int32_t d = 0;
      visit(u, [&](int64_t) { ++d; });
      g->offsets[u + 1] = d;
    }
  });
  for (int64_t u = 0; u < V; ++u) g->offsets[u + 1] += g->offsets[u];
  g->indices.resize((size_t)g->offsets[V]);
  parallel_rows(V, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      int32_t k = g->offsets[u];
      visit(u, [&](int64_t w) { g->indices[k++] = (int32_t)w; });
    }
  });
  return true;
}

// LSD radix sort of 64-bit keys on the low `bits` bits, 8 bits per pass.
static void radix_sort(std::vector<uint64_t>& a, int bits) {
  std::vector<uint64_t> tmp(a.size());
  for (int shift = 0; shift < bits; shift += 8) {
    size_t cnt[257] = {0};
    for (uint64_t x : a) ++cnt[((x >> shift) & 0xFF) + 1];
    for (int i = 0; i < 256; ++i) cnt[i + 1] += cnt[i];
    for (uint64_t x : a) tmp[cnt[(x >> shift) & 0xFF]++] = x;
    a.swap(tmp);
  }
}

bool rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, Csr* g) {
  if (scale < 1 || scale > 30 || edge_factor < 1) return false;
  const int64_t V = 1ll << scale;
  const int64_t M = (int64_t)edge_factor * V;
  std::vector<uint64_t> keys((size_t)(2 * M));
  const double ab = a + b, abc = a + b + c;
  parallel_rows(M, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      uint64_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        const uint64_t h = draw(seed, 3 + (uint64_t)l, (uint64_t)i);
        const double x = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        const uint64_t ub = x >= ab, vb = (x >= a && x < ab) || x >= abc;
        u = (u << 1) | ub;
        v = (v << 1) | vb;
      }
      if (u == v) {
        keys[2 * i] = keys[2 * i + 1] = ~0ull;  // self loop: dropped below
      } else {
        keys[2 * i] = (u << 32) | v;
        keys[2 * i + 1] = (v << 32) | u;
      }
    }
  });
  keys.erase(std::remove(keys.begin(), keys.end(), ~0ull), keys.end());
  radix_sort(keys, 32 + scale);
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  g->nv = (int32_t)V;
  g->offsets.assign(V + 1, 0);
  for (uint64_t k : keys) ++g->offsets[(k >> 32) + 1];
  for (int64_t u = 0; u < V; ++u) g->offsets[u + 1] += g->offsets[u];
  g->indices.resize(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) g->indices[i] = (int32_t)(keys[i] & 0xFFFFFFFFu);
  return true;
}

// A random source with at least one neighbour (R-MAT leaves many vertices
// isolated); deterministic in (seed, k).
int32_t pick_source(const Csr& g, uint64_t seed, uint64_t k) {
  for (uint64_t t = 0; t < 1000000; ++t) {
    const int32_t v = (int32_t)(draw(seed, 97, k * 1000003ull + t) % (uint64_t)g.nv);
    if (g.offsets[v + 1] > g.offsets[v]) return v;
  }
  return 0;
}

}  // namespace gcqg
