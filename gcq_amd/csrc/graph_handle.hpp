// graph_handle.hpp — the gcq_graph handle of the C ABI: a host CSR and its
// lazily uploaded HBM copy, plus the single-traversal context of bfs.hip.
#pragma once
#include "frontier.hpp"
#include "graph.hpp"

struct BfsCtx {
  gcqf::FrontierSet fs;
  int32_t* level = nullptr;
  uint32_t* vis = nullptr;  // visited bitmap (GCQ_BFS_BITMAP)
  void* ctl = nullptr;
  void* hub = nullptr;      // two gcqf::HubPool, by level parity (persistent modes)
  int32_t max_deg = -1;     // the graph's largest degree, -1 until computed
  uint32_t* pinned = nullptr;  // host-visible counters for the host-loop mode
};

struct gcq_graph {
  gcqg::Csr host;
  int32_t* d_off = nullptr;
  int32_t* d_idx = nullptr;
  hipStream_t stream = nullptr;
  BfsCtx ctx;
};

namespace gcqg {
// Copy the CSR to HBM on first use (graphs can be built and inspected on a
// host without a GPU).  Returns a gcq status.
int ensure_uploaded(gcq_graph* g);
}  // namespace gcqg
