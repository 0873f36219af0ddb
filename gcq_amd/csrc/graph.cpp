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
// load_edge_file(path, format, undirected): Matrix Market (.mtx) and SNAP
//   edge-list files with the reference loaders' rules (load_mtx / load_snap,
//   bfs_bench.cpp:202-243), parsed from one read of the file and assembled
//   by a counting sort on the source vertex plus a per-row sort and dedup.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
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

// ------------------------------------------------------------------ edge-list files
namespace {

// One line of text: [b, e).
struct Line {
  const char* b;
  const char* e;
};

// Parse a signed decimal integer as `istream >> long long` does: skip
// leading blanks, optional sign, at least one digit, stop at the first
// non-digit.  False on no digits or overflow.
bool parse_ll(const char*& p, const char* e, long long* out) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  if (p >= e || *p < '0' || *p > '9') return false;
  unsigned long long v = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    const unsigned d = (unsigned)(*p++ - '0');
    if (v > (9223372036854775807ull - d) / 10ull) return false;  // out of range for long long
    v = v * 10ull + d;
  }
  *out = neg ? -(long long)v : (long long)v;
  return true;
}

bool read_file(const char* path, std::string* buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char tmp[1 << 16];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf->append(tmp, n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

template <class F>
void for_lines(const std::string& buf, F f) {
  const char* p = buf.data();
  const char* end = p + buf.size();
  while (p < end) {
    const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    const char* le = nl ? nl : end;
    if (!f(Line{p, le})) return;
    p = nl ? nl + 1 : end;
  }
}

// Sorted, deduplicated CSR over V vertices from (u, v) pairs; pairs with an
// endpoint outside [0, V) are dropped (csr_from_edges, bfs_bench.cpp:185-200).
void csr_from_pairs(int64_t V, const std::vector<int64_t>& us, const std::vector<int64_t>& vs, Csr* g) {
  g->nv = (int32_t)V;
  std::vector<int64_t> cnt(V + 1, 0);
  const size_t m = us.size();
  for (size_t i = 0; i < m; ++i)
    if (us[i] >= 0 && us[i] < V && vs[i] >= 0 && vs[i] < V) ++cnt[us[i] + 1];
  for (int64_t v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
  std::vector<int32_t> adj(cnt[V]);
  std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
  for (size_t i = 0; i < m; ++i)
    if (us[i] >= 0 && us[i] < V && vs[i] >= 0 && vs[i] < V) adj[cur[us[i]]++] = (int32_t)vs[i];
  g->offsets.assign(V + 1, 0);
  g->indices.clear();
  g->indices.reserve(adj.size());
  for (int64_t v = 0; v < V; ++v) {
    int32_t* b = adj.data() + cnt[v];
    int32_t* e = adj.data() + cnt[v + 1];
    std::sort(b, e);
    e = std::unique(b, e);
    g->indices.insert(g->indices.end(), b, e);
    g->offsets[v + 1] = (int32_t)g->indices.size();
  }
}

}  // namespace

int load_edge_file(const char* path, int format, bool undirected, Csr* g, std::string* err) {
  std::string buf;
  if (!read_file(path, &buf)) {
    *err = std::string("cannot read ") + path;
    return -1;
  }
  if (format == kFormatAuto) {
    const size_t n = std::strlen(path);
    format = (n >= 4 && std::strcmp(path + n - 4, ".mtx") == 0) ? kFormatMtx : kFormatSnap;
  }
  std::vector<int64_t> us, vs;
  auto add = [&](int64_t u, int64_t v) {
    us.push_back(u);
    vs.push_back(v);
    if (undirected) {
      us.push_back(v);
      vs.push_back(u);
    }
  };
  int64_t V = 0;
  if (format == kFormatMtx) {
    // Banner, comment lines, "rows cols nnz", then nnz "i j [value]" lines
    // (1-based); self loops and out-of-range entries are counted but not
    // kept; V = max(rows, cols).
    if (buf.compare(0, 14, "%%MatrixMarket") != 0) {
      *err = std::string(path) + ": no %%MatrixMarket banner";
      return -1;
    }
    long long nr = -1, nc = -1, nz = -1, cnt = 0;
    bool first = true, sized = false, bad_size = false;
    for_lines(buf, [&](Line l) {
      if (first) {
        first = false;
        return true;  // the banner
      }
      if (l.b == l.e || *l.b == '%') return true;
      const char* p = l.b;
      if (!sized) {
        sized = true;
        bad_size = !(parse_ll(p, l.e, &nr) && parse_ll(p, l.e, &nc) && parse_ll(p, l.e, &nz)) || nr < 0 || nc < 0 ||
                   nz < 0 || nr > 2147483647ll || nc > 2147483647ll;
        return !bad_size && nz > 0;
      }
      long long i, j;
      if (!(parse_ll(p, l.e, &i) && parse_ll(p, l.e, &j))) return true;
      --i;
      --j;
      ++cnt;
      if (i != j && i >= 0 && j >= 0 && i < nr && j < nc) add(i, j);
      return cnt < nz;
    });
    if (!sized || bad_size) {
      *err = std::string(path) + ": bad or missing size line";
      return -1;
    }
    V = std::max(nr, nc);
  } else if (format == kFormatSnap) {
    // "u v" lines, '#' / '%' comments; self loops dropped; ids shifted down
    // by one when the smallest id is 1; V = largest id + 1.
    long long mn = 1ll << 60, mx = -1;
    for_lines(buf, [&](Line l) {
      if (l.b == l.e || *l.b == '#' || *l.b == '%') return true;
      const char* p = l.b;
      long long u, v;
      if (!(parse_ll(p, l.e, &u) && parse_ll(p, l.e, &v)) || u == v) return true;
      mn = std::min(mn, std::min(u, v));
      mx = std::max(mx, std::max(u, v));
      add(u, v);
      return true;
    });
    if (mn == 1) {
      for (auto& x : us) --x;
      for (auto& x : vs) --x;
      --mx;
    }
    V = mx + 1;
  } else {
    *err = "unknown graph file format";
    return -1;
  }
  if (V <= 0 || V > 2147483646ll) {
    *err = std::string(path) + ": no vertices (or more than 2^31 - 2)";
    return -1;
  }
  csr_from_pairs(V, us, vs, g);
  if ((int64_t)g->indices.size() >= (1ll << 31)) {
    *err = std::string(path) + ": more than 2^31 - 1 edges";
    return -1;
  }
  return 0;
}

}  // namespace gcqg
