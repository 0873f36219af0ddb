// batch.hip — batched BFS: many traversals over a set of resident graphs in
// one persistent kernel, pulling work through one shared frontier queue.
//
// No counterpart in the reference (bfs_bench.cpp runs one query per
// process).  Semantics per query are those of bfs_bench.cpp:327-633: a
// vertex discovered while expanding level L gets depth L+1 through a CAS on
// -1, and each query's depth array is bit-exact with a plain run.
//
// Design (MI355X):
//  * Work items are (query slot, vertex) packed in 32 bits: the vertex in
//    the low `vbits` bits (vbits = bits of the largest resident V), the slot
//    above it, below 2^31 so no queue sentinel can occur.  One queue
//    operation per item, carried in the ring slot itself.
//  * All queries advance level-synchronously through one grid barrier per
//    level, so the barrier and queue latency of a level is paid once for the
//    whole batch instead of once per query; items of different queries share
//    every wave (per-lane graph pointers, wave-cooperative expansion).
//  * A query is finished when a level discovered nothing for it: every
//    discovery stores L+1 into the query's `alive` word (its own 128-B line,
//    a plain store, no RMW) and block 0 checks the words after each barrier.
//  * The launch sequence (depth fill, queue reset, seeding, traversal,
//    per-query statistics) can be captured into a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bfs_batch.hpp"
#include "graph_handle.hpp"

namespace bk {

struct Ctl {
  Barrier bar;
  uint32_t next[3], p0[29];
  uint32_t err, p1[31];
  int32_t levels, p2[31];
};

template <class F>
__global__ void k_batch(Frontiers<F> fq, Dev d, Ctl* ctl, HubPool* hub) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  uint32_t* st = stage[(threadIdx.x / 64) & 15];
  uint64_t prev_end[2] = {kNoHead, kNoHead};  // per frontier: its end when last drained
  int32_t L = 0;
  for (;; ++L) {
    if (F::kQueues == 3) fq.idle(L).recycle(tid, nthreads, tid);
    // The frontier drained at the last barrier is this level's output: its
    // Head goes back to that level's end (the FAA pops overshoot it), off
    // the barrier's path; the next barrier orders the store before the level
    // that reads it.
    if constexpr (F::kQueues == 2)
      if (L > 0 && blockIdx.x == 0 && threadIdx.x == 0) fq.in(L - 1).drained(prev_end[(L - 1) & 1]);
    uint64_t end = 0;
    HubPool* const hp = &hub[L & 1];
    const uint32_t dw = wave_sum32(expand<F, false>(fq.in(L), fq.out(L), d, L, tid, st,
                                                    F::kHeadFromDrain ? prev_end[L & 1] : kNoHead, nthreads / 64, &end,
                                                    hp));
    prev_end[L & 1] = end;
    if (lane_id() == 0 && dw) gcq::faa_rlx(&ctl->next[L % 3], dw);
    auto level_end = [&]() {
      if (F::kQueues == 3) fq.idle(L).rewind();
      hub_reset(&hub[(L + 1) & 1]);
    };
    uint32_t hubs_posted = 0;
    if (!grid_barrier(&ctl->bar, gridDim.x, [&]() -> uint32_t {
          if (gcq::ld_rlx(&hp->word) != 0ull) return 1u;  // the level ends after its hub scan
          level_end();
          return 0u;
        }, &hubs_posted))
      break;
    // Drop this CU's and XCD's cached lines so the cached visited probes of
    // the next level see at least every bit set up to this barrier.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (hubs_posted) {
      const uint32_t dh = wave_sum32(hubs<F, false>(fq.out(L), d, hp, L, tid, (uint32_t)(nthreads / 64), st));
      if (lane_id() == 0 && dh) gcq::faa_rlx(&ctl->next[L % 3], dh);
      if (!grid_barrier(&ctl->bar, gridDim.x, level_end)) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const uint32_t n = gcq::ld_rlx(&ctl->next[L % 3]);
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) gcq::st_rlx(&ctl->next[(L + 2) % 3], 0u);
      // Queries with nothing discovered at level L are finished.  alive only
      // grows, so a block already at level L+1 cannot hide a finish.
      for (uint32_t s = threadIdx.x; s < d.nslots; s += blockDim.x) {
        const int32_t start = (int32_t)gcq::ld_rlx(&d.slots[s].start);
        if (d.done[s] < 0 && start <= L && (int32_t)gcq::ld_rlx(&d.alive[s * kAliveStride]) < L + 1)
          d.done[s] = L + 1 - start;
      }
    }
    if (n == 0) break;
  }
  if (tid == 0) ctl->levels = L + 1;
}

// Multi-source groups (bfs_batch.hpp): the same level loop over group items.
template <class F>
__global__ void k_batch_ms(Frontiers<F> fq, MsDev d, Ctl* ctl, HubPool* hub) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  uint32_t* st = stage[(threadIdx.x / 64) & 15];
  uint64_t prev_end[2] = {kNoHead, kNoHead};
  int32_t L = 0;
  for (;; ++L) {
    if (F::kQueues == 3) fq.idle(L).recycle(tid, nthreads, tid);
    if constexpr (F::kQueues == 2)
      if (L > 0 && blockIdx.x == 0 && threadIdx.x == 0) fq.in(L - 1).drained(prev_end[(L - 1) & 1]);
    uint64_t end = 0;
    HubPool* const hp = &hub[L & 1];
    const uint32_t dw = wave_sum32(ms_expand<F>(fq.in(L), fq.out(L), d, L, tid, st,
                                                F::kHeadFromDrain ? prev_end[L & 1] : kNoHead, nthreads / 64, &end,
                                                hp));
    prev_end[L & 1] = end;
    if (lane_id() == 0 && dw) gcq::faa_rlx(&ctl->next[L % 3], dw);
    auto level_end = [&]() {
      if (F::kQueues == 3) fq.idle(L).rewind();
      hub_reset(&hub[(L + 1) & 1]);
    };
    uint32_t hubs_posted = 0;
    if (!grid_barrier(&ctl->bar, gridDim.x, [&]() -> uint32_t {
          if (gcq::ld_rlx(&hp->word) != 0ull) return 1u;  // the level ends after its hub scan
          level_end();
          return 0u;
        }, &hubs_posted))
      break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (hubs_posted) {
      const uint32_t dh = wave_sum32(ms_hubs(fq.out(L), d, hp, L, tid, (uint32_t)(nthreads / 64), st));
      if (lane_id() == 0 && dh) gcq::faa_rlx(&ctl->next[L % 3], dh);
      if (!grid_barrier(&ctl->bar, gridDim.x, level_end)) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const uint32_t n = gcq::ld_rlx(&ctl->next[L % 3]);
    if (blockIdx.x == 0 && threadIdx.x == 0) gcq::st_rlx(&ctl->next[(L + 2) % 3], 0u);
    if (n == 0) break;
  }
  if (tid == 0) ctl->levels = L + 1;
}

// Level-0 mask of each group: query i sets its bit at its source, and the
// fetch-or that makes the word non-zero queues (group, source).
template <class F>
__global__ void k_seed_ms(F q, MsDev d, const uint32_t* qinfo, uint32_t nq) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = tid; i < nq; i += gridDim.x * blockDim.x) {
    const uint32_t g = qinfo[3 * i], bit = qinfo[3 * i + 1], src = qinfo[3 * i + 2];
    uint64_t* in0 = (uint64_t*)d.groups[g].nxt[1];  // read at level 0 as nxt[(0 + 1) & 1]
    if (gcq::or_rlx(&in0[src], 1ull << bit) == 0ull && !q.push(tid, (g << d.vbits) | src)) gcq::or_rlx(d.err, 1u);
  }
}

// Level-0 frontier: one item per query.  The grid has at most as many
// threads as the traversal (queue thread ids must stay below that).
template <class F>
__global__ void k_seed(F q, Dev d) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t s = tid; s < d.nslots; s += gridDim.x * blockDim.x) {
    const Slot& sl = d.slots[s];
    ((int32_t*)sl.depth)[sl.src] = 0;
    if (!q.push(tid, (s << d.vbits) | sl.src)) gcq::or_rlx(d.err, 1u);
  }
}

// Per query: reached vertices, the degree sum over them (the reference's
// edges_scanned: every reached vertex is dequeued once and its adjacency
// scanned, bfs_bench.cpp:372-377) and its levels (largest depth + 1, as the
// reference counts them, bfs_bench.cpp:622-631).
__global__ void k_stats(Dev d, uint64_t* stats) {
  const uint32_t s = blockIdx.y;
  const Slot& sl = d.slots[s];
  const int32_t* off = (const int32_t*)sl.off;
  const int32_t* dep = (const int32_t*)sl.depth;
  uint64_t r = 0, e = 0, m = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < sl.nv; v += gridDim.x * blockDim.x) {
    const int32_t dv = dep[v];
    if (dv >= 0) {
      ++r;
      e += (uint64_t)(off[v + 1] - off[v]);
      m = (uint64_t)dv + 1 > m ? (uint64_t)dv + 1 : m;
    }
  }
  r = gcq::wave_sum(r);
  e = gcq::wave_sum(e);
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = shfl64(m, (int)(lane_id() ^ (uint32_t)o));
    m = t > m ? t : m;
  }
  if (lane_id() == 0) {
    if (r) gcq::faa_rlx(&stats[3 * s], r);
    if (e) gcq::faa_rlx(&stats[3 * s + 1], e);
    if (m) gcq::max_rlx(&stats[3 * s + 2], m);
  }
}

}  // namespace bk

struct gcq_batch {
  std::vector<gcq_graph*> graphs;
  hipStream_t stream = nullptr;
  uint32_t vbits = 1;
  gcqf::FrontierSet fs;
  int32_t* depth = nullptr;
  uint64_t depth_cap = 0;
  uint32_t* visited = nullptr;  // per-query bitmaps, each on its own 128-B lines
  uint64_t visited_cap = 0;     // words
  uint64_t visited_used = 0;    // words the current batch uses (a multiple of 32)
  uint32_t slot_cap = 0;
  bk::Slot* d_slots = nullptr;
  bk::Slot* h_slots = nullptr;  // pinned
  uint32_t* alive = nullptr;
  int32_t* done = nullptr;
  uint64_t* stats = nullptr;
  bk::Ctl* ctl = nullptr;
  bk::HubPool* hubs = nullptr;  // device: the two level-parity pools of the multi-source kernel
  // pinned read-back
  int32_t* h_done = nullptr;
  uint64_t* h_stats = nullptr;
  bk::Ctl* h_ctl = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;
  std::vector<uint64_t> depth_at;  // per slot offset into depth
  // multi-source groups (gcq_batch_config.multi_source)
  bk::Group* ms_groups = nullptr;   // device, ms_group_cap entries
  uint64_t* ms_dtab = nullptr;      // device, 64 depth-array pointers per group
  uint32_t* ms_qinfo = nullptr;     // device, (group, bit, source) per query
  uint32_t ms_group_cap = 0, ms_q_cap = 0;
  uint64_t* ms_bits = nullptr;      // device: seen, nxt[0], nxt[1] of every group
  uint64_t ms_bits_cap = 0, ms_bits_used = 0;  // words
  uint32_t ms_ngroups = 0;
  std::vector<bk::Group> h_groups;  // host staging (kept alive across the copies)
  std::vector<uint64_t> h_dtab;
  std::vector<uint32_t> h_qinfo;
};

namespace {

void free_ms(gcq_batch* b) {
  for (void* p : {(void*)b->ms_groups, (void*)b->ms_dtab, (void*)b->ms_qinfo, (void*)b->ms_bits})
    if (p) (void)hipFree(p);
  b->ms_groups = nullptr;
  b->ms_dtab = nullptr;
  b->ms_qinfo = nullptr;
  b->ms_bits = nullptr;
  b->ms_group_cap = b->ms_q_cap = 0;
  b->ms_bits_cap = 0;
}

void free_slots(gcq_batch* b) {
  for (void* p : {(void*)b->d_slots, (void*)b->alive, (void*)b->done, (void*)b->stats})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->h_slots, (void*)b->h_done, (void*)b->h_stats})
    if (p) (void)hipHostFree(p);
  b->d_slots = b->h_slots = nullptr;
  b->alive = nullptr;
  b->done = b->h_done = nullptr;
  b->stats = b->h_stats = nullptr;
  b->slot_cap = 0;
}

void drop_graph(gcq_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key.clear();
}

int ensure_buffers(gcq_batch* b, uint32_t nq, uint64_t total_v, uint64_t visited_words) {
  if (visited_words > b->visited_cap) {
    drop_graph(b);
    if (b->visited) (void)hipFree(b->visited);
    b->visited = nullptr;
    b->visited_cap = 0;
    GCQ_HIP(hipMalloc(&b->visited, 4ull * visited_words));
    b->visited_cap = visited_words;
  }
  if (total_v > b->depth_cap) {
    drop_graph(b);
    if (b->depth) (void)hipFree(b->depth);
    b->depth = nullptr;
    b->depth_cap = 0;
    GCQ_HIP(hipMalloc(&b->depth, 4ull * total_v));
    b->depth_cap = total_v;
  }
  if (nq > b->slot_cap) {
    drop_graph(b);
    free_slots(b);
    GCQ_HIP(hipMalloc(&b->d_slots, sizeof(bk::Slot) * nq));
    GCQ_HIP(hipMalloc(&b->alive, 4ull * bk::kAliveStride * nq));
    GCQ_HIP(hipMalloc(&b->done, 4ull * nq));
    GCQ_HIP(hipMalloc(&b->stats, 24ull * nq));
    GCQ_HIP(hipHostMalloc((void**)&b->h_slots, sizeof(bk::Slot) * nq));
    GCQ_HIP(hipHostMalloc((void**)&b->h_done, 4ull * nq));
    GCQ_HIP(hipHostMalloc((void**)&b->h_stats, 24ull * nq));
    b->slot_cap = nq;
  }
  if (!b->ctl) {
    GCQ_HIP(hipMalloc(&b->ctl, sizeof(bk::Ctl)));
    GCQ_HIP(hipMalloc(&b->hubs, 2 * sizeof(bk::HubPool)));
    GCQ_HIP(hipHostMalloc((void**)&b->h_ctl, sizeof(bk::Ctl)));
  }
  return GCQ_OK;
}

__global__ void k_fill_done(int32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = -1;
}

// Stream-ordered launch sequence of one batch (capturable: kernels only).
template <class F>
int enqueue_batch(gcq_batch* b, const gcqf::Frontiers<F>& fq_in, const bk::Dev& d, uint64_t total_v, uint32_t max_nv,
                  const gcq_batch_config* cfg) {
  hipStream_t s = b->stream;
  // Both batched kernels drain their frontiers by counted pops only: G-LFQ
  // pushes store first-lap entries.
  gcqf::Frontiers<F> fq = fq_in;
  if constexpr (std::is_same_v<F, gcqf::FLfq>)
    for (auto& x : fq.q) x.counted_push = 1u;
  const uint32_t nq = d.nslots;
  hipLaunchKernelGGL(gcqf::kf_fill_i32, dim3(gcqf::grid_of(total_v)), dim3(256), 0, s, b->depth, total_v, -1);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(b->visited_used / 2)), dim3(256), 0, s,
                     (uint64_t*)b->visited, b->visited_used / 2);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(2ull * bk::kAliveStride * nq)), dim3(256), 0, s,
                     (uint64_t*)b->alive, (uint64_t)bk::kAliveStride * nq / 2);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(3ull * nq)), dim3(256), 0, s, b->stats, 3ull * nq);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, s, (uint64_t*)b->ctl,
                     (uint64_t)(sizeof(bk::Ctl) / 8));
  hipLaunchKernelGGL(k_fill_done, dim3((nq + 255) / 256), dim3(256), 0, s, b->done, nq);
  b->fs.reset(s);
  const uint32_t seed_blocks = std::min((nq + 63) / 64, cfg->threads / 64);
  if (cfg->multi_source) {
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(b->ms_bits_used)), dim3(256), 0, s, b->ms_bits,
                       b->ms_bits_used);
    const bk::MsDev md{b->ms_groups, &b->ctl->err, b->ms_ngroups, b->vbits, cfg->pop_width, cfg->flush_at};
    hipLaunchKernelGGL(bk::k_seed_ms<F>, dim3(seed_blocks), dim3(64), 0, s, fq.q[0], md, (const uint32_t*)b->ms_qinfo,
                       nq);
    for (int p = 0; p < 2; ++p)
      hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, s, (uint64_t*)(b->hubs + p), bk::kHubHeaderWords);
    hipLaunchKernelGGL(bk::k_batch_ms<F>, dim3(cfg->threads / cfg->block), dim3(cfg->block), 0, s, fq, md, b->ctl,
                       b->hubs);
  } else {
    hipLaunchKernelGGL(bk::k_seed<F>, dim3(seed_blocks), dim3(64), 0, s, fq.q[0], d);
    for (int p = 0; p < 2; ++p)
      hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, s, (uint64_t*)(b->hubs + p), bk::kHubHeaderWords);
    hipLaunchKernelGGL(bk::k_batch<F>, dim3(cfg->threads / cfg->block), dim3(cfg->block), 0, s, fq, d, b->ctl,
                       b->hubs);
  }
  const uint32_t bx = std::max<uint32_t>(1, std::min<uint32_t>(64, (max_nv + 255) / 256));
  hipLaunchKernelGGL(bk::k_stats, dim3(bx, nq), dim3(256), 0, s, d, b->stats);
  GCQ_HIP(hipGetLastError());
  return GCQ_OK;
}

int check_config(const gcq_batch_config* c) {
  GCQ_REQUIRE(c->kind >= 0 && c->kind <= 3, "unknown queue kind %d", c->kind);
  GCQ_REQUIRE(c->block >= 64 && c->block <= 1024 && c->block % 64 == 0, "block must be a multiple of 64 in [64,1024]");
  GCQ_REQUIRE(c->threads >= c->block && c->threads % c->block == 0, "threads must be a positive multiple of block");
  GCQ_REQUIRE(c->pop_width >= 1 && c->pop_width <= 64, "pop_width must be in [1,64]");
  GCQ_REQUIRE(c->flush_at >= 1 && c->flush_at <= 64, "flush_at must be in [1,64]");
  GCQ_REQUIRE(c->capacity_log2 <= 31, "capacity_log2 must be <= 31");
  GCQ_REQUIRE(c->multi_source <= 1, "multi_source must be 0 or 1");
  return GCQ_OK;
}

// Multi-source groups: queries sorted by graph (stably), up to 64 per group;
// each group gets seen / nxt[0] / nxt[1] words for its graph's vertices.
int build_groups(gcq_batch* b, int32_t nq, const int32_t* gids, const uint32_t* srcs,
                 const std::vector<uint64_t>& depth_at, uint64_t* group_v) {
  std::vector<int32_t> order(nq);
  for (int32_t i = 0; i < nq; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return gids[x] < gids[y]; });
  b->h_groups.clear();
  b->h_dtab.clear();
  b->h_qinfo.assign(3ull * nq, 0u);
  uint64_t words = 0;
  *group_v = 0;
  for (int32_t k = 0; k < nq;) {
    const int32_t gi = gids[order[k]];
    const gcq_graph* gr = b->graphs[gi];
    const uint64_t nv = (uint64_t)gr->host.nv;
    bk::Group G{};
    G.off = (uint64_t)gr->d_off;
    G.idx = (uint64_t)gr->d_idx;
    G.start = 0;  // every member starts at level 0
    G.fresh = 0;
    G.seen = words;  // offsets for now: turned into addresses once ms_bits exists
    G.nxt[0] = words + nv;
    G.nxt[1] = words + 2 * nv;
    words += 3 * nv;
    *group_v += nv;
    const uint32_t g = (uint32_t)b->h_groups.size();
    G.depth = 64ull * g;  // index into ms_dtab, likewise
    uint32_t m = 0;
    for (; k < nq && gids[order[k]] == gi && m < 64; ++k, ++m) {
      const int32_t q = order[k];
      b->h_qinfo[3 * q] = g;
      b->h_qinfo[3 * q + 1] = m;
      b->h_qinfo[3 * q + 2] = srcs[q];
      b->h_dtab.push_back((uint64_t)(b->depth + depth_at[q]));
    }
    while (b->h_dtab.size() % 64) b->h_dtab.push_back(0);
    b->h_groups.push_back(G);
  }
  const uint32_t ng = (uint32_t)b->h_groups.size();
  if (ng > b->ms_group_cap || (uint32_t)nq > b->ms_q_cap || words > b->ms_bits_cap) {
    drop_graph(b);
    free_ms(b);
    GCQ_HIP(hipMalloc(&b->ms_groups, sizeof(bk::Group) * ng));
    GCQ_HIP(hipMalloc(&b->ms_dtab, 8ull * 64 * ng));
    GCQ_HIP(hipMalloc(&b->ms_qinfo, 12ull * nq));
    GCQ_HIP(hipMalloc(&b->ms_bits, 8ull * words));
    b->ms_group_cap = ng;
    b->ms_q_cap = (uint32_t)nq;
    b->ms_bits_cap = words;
  }
  for (bk::Group& G : b->h_groups) {
    G.seen = (uint64_t)(b->ms_bits + G.seen);
    G.nxt[0] = (uint64_t)(b->ms_bits + G.nxt[0]);
    G.nxt[1] = (uint64_t)(b->ms_bits + G.nxt[1]);
    G.depth = (uint64_t)(b->ms_dtab + G.depth);
  }
  b->ms_ngroups = ng;
  b->ms_bits_used = words;
  GCQ_HIP(hipMemcpyAsync(b->ms_groups, b->h_groups.data(), sizeof(bk::Group) * ng, hipMemcpyHostToDevice, b->stream));
  GCQ_HIP(hipMemcpyAsync(b->ms_dtab, b->h_dtab.data(), 8ull * b->h_dtab.size(), hipMemcpyHostToDevice, b->stream));
  GCQ_HIP(hipMemcpyAsync(b->ms_qinfo, b->h_qinfo.data(), 12ull * nq, hipMemcpyHostToDevice, b->stream));
  GCQ_HIP(hipStreamSynchronize(b->stream));  // the staging vectors may change on the next call
  return GCQ_OK;
}

int run_batch(gcq_batch* b, const gcq_batch_config* cfg, int32_t nq, const int32_t* gids, const int32_t* srcs,
              int32_t* const* depth_out, gcq_query_result* per_query, gcq_batch_result* res) {
  int rc = check_config(cfg);
  if (rc) return rc;
  GCQ_REQUIRE(nq > 0 && gids && srcs && res, "null argument or empty batch");
  GCQ_REQUIRE((uint64_t)nq <= (1ull << (31 - b->vbits)), "batch of %d queries exceeds the %llu slots of %u-bit vertices",
              nq, 1ull << (31 - b->vbits), b->vbits);
  std::memset(res, 0, sizeof *res);
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t total_v = 0, total_w = 0;
  uint32_t max_nv = 0;
  std::vector<uint64_t> at(nq), wat(nq);
  for (int32_t i = 0; i < nq; ++i) {
    GCQ_REQUIRE(gids[i] >= 0 && gids[i] < (int32_t)b->graphs.size(), "query %d: bad graph id %d", i, gids[i]);
    gcq_graph* g = b->graphs[gids[i]];
    rc = gcqg::ensure_uploaded(g);
    if (rc) return rc;
    at[i] = total_v;
    total_v += (uint64_t)g->host.nv;
    wat[i] = total_w;
    total_w += ((uint64_t)g->host.nv + 1023) / 1024 * 32;  // whole 128-B lines
    max_nv = std::max<uint32_t>(max_nv, (uint32_t)g->host.nv);
  }
  rc = ensure_buffers(b, (uint32_t)nq, total_v, total_w);
  b->visited_used = total_w;
  if (rc) return rc;
  uint64_t need = total_v;  // one item per (query, vertex) at most per level
  if (cfg->multi_source) {
    std::vector<uint32_t> src_ok(nq);
    for (int32_t i = 0; i < nq; ++i) {
      const int32_t nv = b->graphs[gids[i]]->host.nv;
      src_ok[i] = (srcs[i] >= 0 && srcs[i] < nv) ? (uint32_t)srcs[i] : 0u;  // bfs_bench.cpp:555
    }
    rc = build_groups(b, nq, gids, src_ok.data(), at, &need);  // one item per (group, vertex)
    if (rc) return rc;
  }
  uint64_t cap_items = cfg->capacity_log2 ? (1ull << cfg->capacity_log2) : need;
  GCQ_REQUIRE(cap_items >= need, "capacity 2^%u is below the %llu items a level may hold", cfg->capacity_log2,
              (unsigned long long)need);
  const void* fs_before = b->fs.q[0];
  rc = b->fs.ensure(cfg->kind, cap_items, cfg->threads);
  if (rc) return rc;
  if (b->fs.q[0] != fs_before) drop_graph(b);
  for (int32_t i = 0; i < nq; ++i) {
    gcq_graph* g = b->graphs[gids[i]];
    bk::Slot s{};
    s.off = (uint64_t)g->d_off;
    s.idx = (uint64_t)g->d_idx;
    s.depth = (uint64_t)(b->depth + at[i]);
    s.visited = (uint64_t)(b->visited + wat[i]);
    s.start = 0;
    s.src = (srcs[i] >= 0 && srcs[i] < g->host.nv) ? (uint32_t)srcs[i] : 0u;  // bfs_bench.cpp:555
    s.nv = (uint32_t)g->host.nv;
    b->h_slots[i] = s;
  }
  GCQ_HIP(hipMemcpyAsync(b->d_slots, b->h_slots, sizeof(bk::Slot) * nq, hipMemcpyHostToDevice, b->stream));
  bk::Dev d{b->d_slots, b->alive, b->done, &b->ctl->err, (uint32_t)nq, b->vbits, cfg->pop_width, cfg->flush_at};

  // Every block of the persistent kernel must be resident.
  int dev = 0, cus = 0, per_cu = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  rc = b->fs.dispatch([&](const auto& fq) {
    using F = std::decay_t<decltype(fq.q[0])>;
    if (cfg->multi_source)
      GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bk::k_batch_ms<F>, cfg->block, 0));
    else
      GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bk::k_batch<F>, cfg->block, 0));
    return GCQ_OK;
  });
  if (rc) return rc;
  GCQ_REQUIRE((int64_t)(cfg->threads / cfg->block) <= (int64_t)per_cu * cus,
              "persistent batch needs all %u blocks resident; the device holds %d", cfg->threads / cfg->block,
              per_cu * cus);

  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  GCQ_HIP(hipEventRecord(e0, b->stream));
  if (cfg->use_graph) {
    const std::vector<uint64_t> key = {(uint64_t)cfg->kind, cfg->threads, cfg->block, cfg->pop_width, cfg->flush_at,
                                       cfg->capacity_log2, (uint64_t)nq, total_v, max_nv,
                                       (uint64_t)b->fs.q[0], (uint64_t)b->depth, (uint64_t)b->d_slots,
                                       (uint64_t)b->visited, total_w, cfg->multi_source,
                                       (uint64_t)b->ms_bits, b->ms_bits_used, b->ms_ngroups};
    if (key != b->key || !b->exec) {
      drop_graph(b);
      GCQ_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = b->fs.dispatch([&](const auto& fq) { return enqueue_batch(b, fq, d, total_v, max_nv, cfg); });
      hipGraph_t gr = nullptr;
      const hipError_t ce = hipStreamEndCapture(b->stream, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      GCQ_HIP(ce);
      b->graph = gr;
      GCQ_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->key = key;
    } else {
      res->graph_replayed = 1;
    }
    GCQ_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    rc = b->fs.dispatch([&](const auto& fq) { return enqueue_batch(b, fq, d, total_v, max_nv, cfg); });
    if (rc) return rc;
  }
  GCQ_HIP(hipEventRecord(e1, b->stream));
  GCQ_HIP(hipMemcpyAsync(b->h_done, b->done, 4ull * nq, hipMemcpyDeviceToHost, b->stream));
  GCQ_HIP(hipMemcpyAsync(b->h_stats, b->stats, 24ull * nq, hipMemcpyDeviceToHost, b->stream));
  GCQ_HIP(hipMemcpyAsync(b->h_ctl, b->ctl, sizeof(bk::Ctl), hipMemcpyDeviceToHost, b->stream));
  GCQ_HIP(hipStreamSynchronize(b->stream));
  res->kernel_ms = gcqh::ms_between(e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->levels = b->h_ctl->levels;
  for (int32_t i = 0; i < nq; ++i) {
    res->edges_scanned += b->h_stats[3 * i + 1];
    if (per_query) {
      per_query[i].reached = b->h_stats[3 * i];
      per_query[i].edges_scanned = b->h_stats[3 * i + 1];
      per_query[i].levels = (int32_t)b->h_stats[3 * i + 2];  // largest depth + 1
      per_query[i].status = per_query[i].levels > 0 ? 0 : -1;
    }
  }
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (b->h_ctl->bar.abort) {
    gcqh::set_error("batched BFS aborted: a grid-barrier wait exceeded its spin limit");
    return GCQ_ETIMEOUT;
  }
  if (b->h_ctl->err) {
    gcqh::set_error("batched BFS frontier enqueue failed");
    return GCQ_EFULL;
  }
  if (depth_out) {
    for (int32_t i = 0; i < nq; ++i)
      if (depth_out[i])
        GCQ_HIP(hipMemcpyAsync(depth_out[i], b->depth + at[i], 4ull * b->graphs[gids[i]]->host.nv,
                               hipMemcpyDeviceToHost, b->stream));
    GCQ_HIP(hipStreamSynchronize(b->stream));
  }
  return GCQ_OK;
}

}  // namespace

extern "C" void gcq_batch_config_default(gcq_batch_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->kind = GCQ_GLFQ;
  c->threads = 16384;
  c->block = 256;
  c->pop_width = 64;
  c->flush_at = 64;
  c->capacity_log2 = 0;
  c->use_graph = 1;
}

extern "C" int gcq_batch_create(gcq_graph* const* graphs, int32_t ngraphs, gcq_batch** out) {
  GCQ_REQUIRE(out && graphs && ngraphs > 0, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  gcq_batch* b = new gcq_batch();
  int32_t max_nv = 1;
  for (int32_t i = 0; i < ngraphs; ++i) {
    if (!graphs[i]) {
      delete b;
      GCQ_REQUIRE(false, "graph %d is null", i);
    }
    b->graphs.push_back(graphs[i]);
    max_nv = std::max(max_nv, graphs[i]->host.nv);
  }
  while ((1ll << b->vbits) < (int64_t)max_nv) ++b->vbits;
  if (b->vbits > 30) {
    delete b;
    GCQ_REQUIRE(false, "graphs of %d vertices leave no room for query slots", max_nv);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    gcqh::set_error("hipStreamCreate failed");
    return GCQ_EHIP;
  }
  *out = b;
  return GCQ_OK;
}

extern "C" int gcq_batch_run(gcq_batch* b, const gcq_batch_config* cfg, int32_t nq, const int32_t* graph_ids,
                             const int32_t* sources, int32_t* const* depth_out, gcq_query_result* per_query,
                             gcq_batch_result* res) {
  GCQ_REQUIRE(b && cfg, "null argument");
  return run_batch(b, cfg, nq, graph_ids, sources, depth_out, per_query, res);
}

extern "C" int gcq_batch_autotune(gcq_batch* b, int32_t kind, int32_t nq, const int32_t* graph_ids,
                                  const int32_t* sources, int32_t reps, gcq_batch_config* best,
                                  gcq_tune_trial* trials_out, int32_t max_trials, int32_t* ntrials) {
  GCQ_REQUIRE(b && best && nq > 0, "null argument");
  if (reps < 1) reps = 1;
  int dev = 0, cus = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  uint64_t total_v = 0;
  for (int32_t i = 0; i < nq; ++i) {
    GCQ_REQUIRE(graph_ids[i] >= 0 && graph_ids[i] < (int32_t)b->graphs.size(), "bad graph id");
    total_v += (uint64_t)b->graphs[graph_ids[i]]->host.nv;
  }
  uint32_t min_log2 = 6;
  while ((1ull << min_log2) < total_v) ++min_log2;
  gcq_batch_config cur;
  gcq_batch_config_default(&cur);
  cur.kind = kind;
  cur.use_graph = 0;
  int32_t nt = 0;
  // Median of reps timed runs after one warm run; a failing configuration
  // (e.g. not all blocks resident) is skipped.
  auto measure = [&](const gcq_batch_config& c, double* ms) -> int {
    gcq_batch_result r;
    int rc = run_batch(b, &c, nq, graph_ids, sources, nullptr, nullptr, &r);
    if (rc) return rc;
    std::vector<double> t;
    for (int32_t k = 0; k < reps; ++k) {
      rc = run_batch(b, &c, nq, graph_ids, sources, nullptr, nullptr, &r);
      if (rc) return rc;
      t.push_back(r.total_ms);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    if (trials_out && nt < max_trials) trials_out[nt] = gcq_tune_trial{c, *ms};
    ++nt;
    return GCQ_OK;
  };
  double best_ms = 1e300;
  int rc = measure(cur, &best_ms);
  if (rc) return rc;
  // Coordinate descent, two sweeps over the knobs: the traversal form
  // (multi-source groups or one item per query) first, then the grid.
  for (int sweep = 0; sweep < 2; ++sweep) {
    for (int knob : {5, 0, 1, 2, 3, 4}) {
      std::vector<gcq_batch_config> cand;
      gcq_batch_config c = cur;
      switch (knob) {
        case 5:
          for (uint32_t ms : {0u, 1u}) {
            c = cur;
            c.multi_source = ms;
            cand.push_back(c);
          }
          break;
        case 0:
          for (uint32_t blk : {128u, 256u, 512u}) {
            c = cur;
            c.block = blk;
            c.threads = std::max(blk, cur.threads / blk * blk);
            cand.push_back(c);
          }
          break;
        case 1:
          for (uint32_t wpc : {2u, 4u, 8u, 16u}) {  // waves per CU
            c = cur;
            c.threads = std::max(c.block, (uint32_t)cus * wpc * 64u / c.block * c.block);
            cand.push_back(c);
          }
          break;
        case 2:
          for (uint32_t pw : {16u, 32u, 64u}) {
            c = cur;
            c.pop_width = pw;
            cand.push_back(c);
          }
          break;
        case 3:
          for (uint32_t fa : {16u, 32u, 64u}) {
            c = cur;
            c.flush_at = fa;
            cand.push_back(c);
          }
          break;
        case 4:
          for (uint32_t cl : {min_log2, min_log2 + 1}) {
            c = cur;
            c.capacity_log2 = cl;
            cand.push_back(c);
          }
          break;
      }
      for (const auto& x : cand) {
        if (std::memcmp(&x, &cur, sizeof x) == 0) continue;
        double ms = 0;
        if (measure(x, &ms) != GCQ_OK) continue;
        if (ms < best_ms) {
          best_ms = ms;
          cur = x;
        }
      }
    }
  }
  cur.use_graph = 1;
  *best = cur;
  if (ntrials) *ntrials = nt;
  return GCQ_OK;
}

extern "C" int gcq_batch_destroy(gcq_batch* b) {
  if (!b) return GCQ_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  b->fs.release();
  free_slots(b);
  free_ms(b);
  if (b->depth) (void)hipFree(b->depth);
  if (b->visited) (void)hipFree(b->visited);
  if (b->ctl) (void)hipFree(b->ctl);
  if (b->hubs) (void)hipFree(b->hubs);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GCQ_OK;
}
