// bfs.cpp — ORACLE (test infrastructure only).  Textbook BFS on an int32 CSR
// for bit-exact depth checks, and a level-synchronous top-down BFS over host
// threads for the CPU baseline.  The reference validates no depths
// (bfs_bench.cpp:778-787 prints reached/levels only); BFS depths are unique
// for a given graph and source, so any correct BFS must match these.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

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

extern "C" long orc_bfs_parallel(int32_t nv, const int32_t* offs, const int32_t* idx, int32_t src, int32_t* depth,
                                 int threads) {
  if (nv <= 0) return 0;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (src < 0 || src >= nv) src = 0;
  std::vector<std::atomic<int32_t>> d(nv);
  for (int32_t i = 0; i < nv; ++i) d[i].store(-1, std::memory_order_relaxed);
  d[src].store(0);
  std::vector<int32_t> cur{src}, next;
  std::vector<std::vector<int32_t>> local(threads);
  long reached = 1;
  for (int32_t level = 0; !cur.empty(); ++level) {
    std::vector<std::thread> ts;
    const size_t chunk = (cur.size() + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
      ts.emplace_back([&, t]() {
        auto& out = local[t];
        out.clear();
        const size_t b = t * chunk, e = std::min(cur.size(), b + chunk);
        for (size_t i = b; i < e; ++i) {
          const int32_t u = cur[i];
          for (int32_t k = offs[u]; k < offs[u + 1]; ++k) {
            const int32_t w = idx[k];
            int32_t exp = -1;
            if (d[w].load(std::memory_order_relaxed) == -1 &&
                d[w].compare_exchange_strong(exp, level + 1, std::memory_order_relaxed))
              out.push_back(w);
          }
        }
      });
    }
    for (auto& x : ts) x.join();
    next.clear();
    for (auto& v : local) next.insert(next.end(), v.begin(), v.end());
    reached += (long)next.size();
    cur.swap(next);
  }
  for (int32_t i = 0; i < nv; ++i) depth[i] = d[i].load(std::memory_order_relaxed);
  return reached;
}
