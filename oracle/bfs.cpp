// bfs.cpp — ORACLE (test infrastructure only).  Textbook BFS on an int32 CSR
// for bit-exact depth checks, and a level-synchronous top-down OpenMP BFS
// for the CPU baseline.  The reference validates no depths
// (bfs_bench.cpp:778-787 prints reached/levels only); BFS depths are unique
// for a given graph and source, so any correct BFS must match these.
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <thread>
#include <vector>

#include <omp.h>

#include "oracle.h"

extern "C" long orc_bfs_serial(int32_t nv, const int32_t* offs, const int32_t* idx, int32_t src, int32_t* depth,
                               uint64_t* edges_scanned) {
  if (nv <= 0) return 0;
  std::fill(depth, depth + nv, -1);
  if (src < 0 || src >= nv) src = 0;  // bfs_bench.cpp:555
  std::vector<int32_t> q;
  q.reserve(1024);
  q.push_back(src);
  depth[src] = 0;
  uint64_t es = 0;
  for (size_t h = 0; h < q.size(); ++h) {
    const int32_t u = q[h];
    es += (uint64_t)(offs[u + 1] - offs[u]);
    for (int32_t e = offs[u]; e < offs[u + 1]; ++e) {
      const int32_t w = idx[e];
      if (depth[w] < 0) {
        depth[w] = depth[u] + 1;
        q.push_back(w);
      }
    }
  }
  if (edges_scanned) *edges_scanned = es;
  return (long)q.size();
}

// Level-synchronous top-down BFS with OpenMP (the CPU baseline): one
// parallel region for the whole traversal; per level the frontier is split
// dynamically over the threads, each claims unvisited neighbours with a CAS
// on the depth word and keeps its discoveries in a private buffer, and the
// buffers are concatenated into the next frontier after a prefix sum of
// their sizes.  Threads are bound as OMP_PROC_BIND says (close in the
// baseline runs).
extern "C" long orc_bfs_parallel(int32_t nv, const int32_t* offs, const int32_t* idx, int32_t src, int32_t* depth,
                                 int threads) {
  if (nv <= 0) return 0;
  if (threads <= 0) threads = omp_get_max_threads();
  if (src < 0 || src >= nv) src = 0;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int32_t i = 0; i < nv; ++i) depth[i] = -1;
  depth[src] = 0;
  std::vector<int32_t> cur((size_t)nv), next((size_t)nv);
  std::vector<size_t> counts((size_t)threads + 1, 0);
  cur[0] = src;
  size_t cur_n = 1, next_n = 0;
  long reached = 1;
  int32_t level = 0;
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    std::vector<int32_t> buf;
    buf.reserve(4096);
    for (;;) {
      const size_t n = cur_n;
      if (n == 0) break;
      buf.clear();
#pragma omp for schedule(dynamic, 64) nowait
      for (size_t i = 0; i < n; ++i) {
        const int32_t u = cur[i];
        for (int32_t k = offs[u]; k < offs[u + 1]; ++k) {
          const int32_t w = idx[k];
          int32_t exp = -1;
          if (__atomic_load_n(&depth[w], __ATOMIC_RELAXED) == -1 &&
              __atomic_compare_exchange_n(&depth[w], &exp, level + 1, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            buf.push_back(w);
        }
      }
      counts[t + 1] = buf.size();
#pragma omp barrier
#pragma omp single
      {
        counts[0] = 0;
        for (int k = 0; k < nt; ++k) counts[k + 1] += counts[k];
        next_n = counts[nt];
      }
      std::copy(buf.begin(), buf.end(), next.begin() + (ptrdiff_t)counts[t]);
#pragma omp barrier
#pragma omp single
      {
        cur.swap(next);
        cur_n = next_n;
        reached += (long)next_n;
        ++level;
      }
    }
  }
  return reached;
}
