// bfs.hip — queue-driven level-synchronous BFS on an int32 CSR in HBM.
//
// Semantics follow the reference harness (bfs_bench.cpp:327-633): two
// frontier queues ping-pong between levels; every dequeued vertex u has its
// neighbours w tested with CAS(level[w], -1, L+1) and the winner enqueues w;
// the traversal ends when a level discovers nothing.  Depths are therefore
// exact BFS depths (bit-exact against a serial BFS).
//
// MI355X design (own):
//  * frontier queues hold vertex ids directly in the ring (G-WFQ / G-LFQ
//    ring-level API, SFQ items, WFQ values + 1): one ring operation per
//    vertex instead of the reference's two-ring indirection plus fence;
//    ring capacity >= V so no vertex is dropped (the reference's 32,768-slot
//    rings drop frontier vertices silently once a level exceeds them,
//    bfs_bench.cpp:38-39, 380);
//  * expansion is wave-cooperative for every degree: the wave prefix-sums the
//    degrees of the (up to 64) vertices it dequeued and strides its 64 lanes
//    over their concatenated adjacency (load balanced; no separate hub
//    kernel, bfs_bench.cpp:434-496);
//  * visited vertices are filtered with a plain load before the CAS;
//  * newly discovered vertices are staged per wave in LDS and enqueued 64 at
//    a time (one aggregated ticket FAA per 64 vertices);
//  * GCQ_BFS_PERSISTENT runs all levels in one launch with a device-side grid
//    barrier (no host round trip per level; the 4900^2 lattice has ~10^4
//    levels).  Every block is co-resident (checked on the host) and every
//    barrier wait has a spin limit that aborts the run instead of hanging.
//    Its barrier carries arrival and the level's discovery count in one
//    64-bit FAA per block, and every frontier kind is drained by counted
//    pops (frontier.hpp): 12.2 us per lattice level instead of 21.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include <type_traits>

#include "graph_handle.hpp"

// ======================================================================== graph handle
namespace bfsk {

using namespace gcqf;

// Control block: the persistent kernel's level barrier, then each contended
// counter on its own 128-B line (next[] counts the host-loop levels).
struct Ctl {
  LevelBarrier lbar;
  uint32_t next[3], p2[29];
  uint32_t err, p3[31];
  uint64_t edges, reached, p4[14];
  int32_t levels, p5[31];
  uint64_t pending, p6[15];  // barrier-free mode: items pushed and not yet fully expanded
};

struct LevelAcc {
  uint64_t edges;
  uint32_t disc;
};

// Drain qin (level L) and enqueue the vertices discovered at depth L+1.
// All 64 lanes of the wave must call it together (threads % 64 == 0).
// kCounted: pop with F::pop_counted against the level's end (persistent
// kernel only: the end holds because every push of the level before was
// complete at its barrier, and the barrier sets Head back afterwards).
// Each neighbour's depth is tested with a cached load before its CAS when
// the wave's batch has more than kProbeEdges edges.  The probe saves an
// atomic per visited neighbour, which pays on long adjacency scans.  It also
// adds a dependent round trip to each 64-edge round, which dominates a batch
// of a few rounds.  Measured: the 4900^2 lattice (about 3 edges per vertex)
// runs 10% faster without the probe; R-MAT scale 20 at 65,536 threads runs
// 18% slower without it.
constexpr int32_t kProbeEdges = 256;

// One pop round per level: every wave of the grid starts a level with one
// pop of all 64 lanes, so that first round takes the tickets [h0, h0 +
// lanes).  When the level's tickets [h0, end) fit in it, no later pop can
// find an item, and a wave leaves after its first batch instead of paying
// one more FAA round trip to learn that.  h0 = kNoHead when the level's
// starting Head is not known.

// kBitmap (GCQ_BFS_BITMAP): a vertex is claimed in the visited bitmap `vis`
// (one bit per vertex, 1,024 vertices per 128-B line) instead of by a CAS on
// its depth word.  Each edge first reads its bit through the caches (bits
// only ever go from 0 to 1, so a stale read can only send a visited vertex
// on to the atomic), a clear bit is claimed by fetch-or, and the one lane
// whose fetch-or set the bit stores the depth with a plain store: nothing
// reads depths during the traversal, and the kernel's end makes them
// visible.  Visited neighbours then cost a cached load instead of an atomic
// on the memory side.
template <bool kCounted, class F, bool kBitmap = false>
__device__ void expand_level(const F& qin, const F& qout, const int32_t* __restrict__ offs,
                             const int32_t* __restrict__ idx, int32_t* level, int32_t L, uint32_t tid, uint32_t* stage,
                             LevelAcc& acc, uint32_t* err, uint64_t h0 = kNoHead, uint64_t grid_lanes = 0,
                             uint64_t* end_out = nullptr, uint32_t* vis = nullptr, HubPool* hp = nullptr) {
  const uint32_t lane = lane_id();
  uint32_t sc = 0;  // staged vertices (wave-uniform)
  uint64_t end = 0;
  bool one_round = false;
  // One-round levels hand out their tickets by grid position, spread over
  // the waves: lane l of wave w takes h0 + l * W + w (W waves), so a level
  // of n items gives each wave about n / W of them, one 64-edge round, and
  // every ticket in [h0, h0 + 64 W) goes to exactly one lane.  Its slot
  // word is read beside the level's end.
  const uint64_t tix = h0 + (uint64_t)lane * (grid_lanes / 64) + tid / 64;
  uint64_t first_word = 0;
  if constexpr (kCounted) {
    if constexpr (F::kHeadFromDrain) {
      if (h0 != kNoHead) first_word = qin.peek(tix);
    }
    end = qin.level_end();
    if (end_out) *end_out = end;
    one_round = h0 != kNoHead && end >= h0 && end - h0 <= grid_lanes;
  }
  for (;;) {
    uint32_t u = 0;
    bool got;
    if constexpr (kCounted) {
      int st;
      if constexpr (F::kHeadFromDrain) {
        // a one-round level takes its tickets by grid position, with no FAA on Head
        st = one_round ? qin.take(tid, tix, end, first_word, &u) : qin.pop_counted(tid, end, &u);
      } else {
        st = qin.pop_counted(tid, end, &u);
      }
      if (__ballot(st >= 0) == 0) break;  // every lane drew a ticket past the level's end
      got = st > 0;
    } else {
      got = qin.pop(tid, &u);
      if (__ballot(got) == 0) break;  // the whole wave saw the level's queue empty
    }
    int32_t beg = 0, deg = 0;
    if (got) {
      beg = offs[u];
      deg = offs[u + 1] - beg;
    }
    acc.edges += (uint64_t)deg;
    if (hp && deg > kHubDeg && hub_post(hp, u, beg, deg, 0ull)) deg = 0;  // the grid scans it
    int32_t incl = deg;  // inclusive prefix of degrees across the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += t;
    }
    const int32_t excl = incl - deg;
    const int32_t total = __shfl(incl, 63, 64);
    const bool probe = total > kProbeEdges;
    auto neighbour = [&](int32_t base) -> uint32_t {  // this lane's edge target in a round, 0 past the end
      const int32_t e = base + (int32_t)lane;
      const int32_t ec = e < total ? e : total - 1;
      int lo = 0, hi = 63;  // owner = first lane whose inclusive prefix exceeds ec
      for (int s = 0; s < 6; ++s) {
        const int mid = (lo + hi) >> 1;
        const int32_t pv = __shfl(incl, mid, 64);
        if (pv > ec) hi = mid;
        else lo = mid + 1;
      }
      const int32_t obeg = __shfl(beg, lo, 64);
      const int32_t oexcl = __shfl(excl, lo, 64);
      return e < total ? (uint32_t)idx[obeg + (e - oexcl)] : 0u;
    };
    // The next round's neighbour indices are loaded before this round's
    // CASes, so the two latencies overlap.
    uint32_t w_next = total > 0 ? neighbour(0) : 0u;
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t e = base + (int32_t)lane;
      const uint32_t w_here = w_next;
      if (base + 64 < total) w_next = neighbour(base + 64);
      bool disc = false;
      uint32_t w = 0;
      if (e < total) {
        w = w_here;
        if constexpr (kBitmap) {
          uint32_t* const vw = vis + (w >> 5);
          const uint32_t vb = 1u << (w & 31u);
          if (!(gcq::ld_cached(vw) & vb)) {
            disc = !(gcq::or_rlx(vw, vb) & vb);
            if (disc) level[w] = L + 1;
          }
        } else if (!probe || gcq::ld_cached(&level[w]) == -1) {
          // A depth changes once per traversal, from -1 to a level, so a cached
          // (possibly stale) read can only send a visited vertex to the CAS.
          int32_t expect = -1;
          disc = gcq::cas_rlx(&level[w], expect, L + 1);
        }
      }
      const uint64_t dm = __ballot(disc);
      if (disc) stage[sc + gcq::rank_in(dm)] = w;
      sc += (uint32_t)__builtin_popcountll(dm);
      acc.disc += disc ? 1u : 0u;
      if (sc >= 64) flush(qout, tid, stage, sc, 64, err);
    }
    if (one_round) break;
  }
  if (sc) flush(qout, tid, stage, sc, sc, err);
}

// The grid's scan of a level's posted vertices (between the level's two
// barriers), as expand_level scans an edge: a probe of the target's depth
// (or visited bit) and a claim.
template <class F, bool kBitmap>
__device__ void hub_level(const F& qout, const int32_t* __restrict__ idx, int32_t* level, int32_t L, uint32_t tid,
                          uint32_t nwaves, uint32_t* stage, LevelAcc& acc, uint32_t* err, uint32_t* vis,
                          HubPool* hp) {
  const uint32_t lane = lane_id();
  uint32_t sc = 0;
  hub_scan(hp, tid, nwaves, [&](const HubEntry* e, int32_t first, int32_t stop) {
    const int32_t* adj = idx + gcq::ld_cached(&e->beg);
    int32_t e0 = first + (int32_t)lane;
    uint32_t w_next = e0 < stop ? (uint32_t)adj[e0] : 0u;
    for (; e0 - (int32_t)lane < stop; e0 += 64) {
      const uint32_t w = w_next;
      if (e0 + 64 < stop) w_next = (uint32_t)adj[e0 + 64];
      bool disc = false;
      if (e0 < stop) {
        if constexpr (kBitmap) {
          uint32_t* const vw = vis + (w >> 5);
          const uint32_t vb = 1u << (w & 31u);
          if (!(gcq::ld_cached(vw) & vb)) {
            disc = !(gcq::or_rlx(vw, vb) & vb);
            if (disc) level[w] = L + 1;
          }
        } else if (gcq::ld_cached(&level[w]) == -1) {
          int32_t expect = -1;
          disc = gcq::cas_rlx(&level[w], expect, L + 1);
        }
      }
      const uint64_t dm = __ballot(disc);
      if (disc) stage[sc + gcq::rank_in(dm)] = w;
      sc += (uint32_t)__builtin_popcountll(dm);
      acc.disc += disc ? 1u : 0u;
      if (sc >= 64) flush(qout, tid, stage, sc, 64, err);
    }
  });
  if (sc) flush(qout, tid, stage, sc, sc, err);
}

template <class F>
__device__ void finish_counts(LevelAcc& acc, uint64_t disc_total, Ctl* c) {
  const uint64_t e = gcq::wave_sum(acc.edges);
  const uint64_t d = gcq::wave_sum(disc_total);
  if (lane_id() == 0) {
    if (e) gcq::faa_rlx(&c->edges, e);
    if (d) gcq::faa_rlx(&c->reached, d);
  }
}

template <class F, bool kBitmap = false>
__global__ void k_persistent(Frontiers<F> fq, const int32_t* __restrict__ offs, const int32_t* __restrict__ idx,
                             int32_t* level, Ctl* c, uint32_t* vis, HubPool* hub) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  uint32_t* st = stage[(threadIdx.x / 64) & 15];
  LevelAcc acc{0, 0};
  uint64_t disc_total = 0;
  uint64_t prev_end[2] = {kNoHead, kNoHead};  // per frontier: its end when last drained
  uint32_t cum = 0;                            // discoveries over all levels so far
  uint32_t seq = 0;                            // level barriers passed
  int32_t L = 0;
  for (;; ++L) {
    if (F::kQueues == 3) fq.idle(L).recycle(tid, nthreads, tid);
    // The frontier drained at the last barrier is this level's output and
    // is read again only at the next level, so its Head is set back to its
    // end here, off the barrier's path; this block's next arrival orders the
    // store before that level.
    if constexpr (F::kCounted && F::kQueues == 2)
      if (L > 0 && blockIdx.x == 0 && threadIdx.x == 0) fq.in(L - 1).drained(prev_end[(L - 1) & 1]);
    acc.disc = 0;
    uint64_t end = 0;
    HubPool* const hp = hub ? &hub[L & 1] : nullptr;
    expand_level<F::kCounted, F, kBitmap>(fq.in(L), fq.out(L), offs, idx, level, L, tid, st, acc, &c->err,
                                          F::kHeadFromDrain ? prev_end[L & 1] : kNoHead, nthreads, &end, vis, hp);
    prev_end[L & 1] = end;
    disc_total += acc.disc;
    // A level that posted long adjacencies ends after the grid's scan of
    // them, at a second barrier; either way the level's last barrier
    // clears the other parity's pool.
    auto level_end = [&]() -> uint32_t {
      uint32_t f = 0;
      if constexpr (F::kQueues == 3) {
        fq.idle(L).rewind();
        f = kBarrierStores;
      }
      if (hub) {
        hub_reset(&hub[(L + 1) & 1]);
        f = kBarrierStores;
      }
      return f;
    };
    uint32_t n = 0;
    bool hubs = false;
    if (!level_barrier(&c->lbar, gridDim.x, seq++, wave_sum32(acc.disc), &n, &cum, [&]() -> uint32_t {
          return hp && gcq::ld_rlx(&hp->word) != 0ull ? kBarrierHubs : level_end();
        }, &hubs))
      break;
    if (hubs) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the posted entries, not stale lines
      acc.disc = 0;
      hub_level<F, kBitmap>(fq.out(L), idx, level, L, tid, (uint32_t)(nthreads / 64), st, acc, &c->err, vis, hp);
      disc_total += acc.disc;
      uint32_t n2 = 0;
      if (!level_barrier(&c->lbar, gridDim.x, seq++, wave_sum32(acc.disc), &n2, &cum, level_end)) break;
      n += n2;
    }
    if (n == 0) break;
  }
  finish_counts<F>(acc, disc_total, c);
  if (tid == 0) c->levels = L + 1;
}

template <class F>
__global__ void k_level(Frontiers<F> fq, const int32_t* __restrict__ offs, const int32_t* __restrict__ idx,
                        int32_t* level, int32_t L, Ctl* c) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) gcq::st_rlx(&c->next[(L + 1) & 1], 0u);  // the next level's counter
  if (F::kQueues == 3) fq.idle(L).recycle(tid, (uint64_t)gridDim.x * blockDim.x, tid);
  LevelAcc acc{0, 0};
  expand_level<false>(fq.in(L), fq.out(L), offs, idx, level, L, tid, stage[(threadIdx.x / 64) & 15], acc, &c->err);
  const uint32_t d = wave_sum32(acc.disc);
  if (lane_id() == 0 && d) gcq::faa_rlx(&c->next[L & 1], d);
  finish_counts<F>(acc, acc.disc, c);
}

// Host-loop mode: rewind the recycled frontier after its level's launch.
template <class F>
__global__ void k_rewind(Frontiers<F> fq, int32_t L) {
  if (threadIdx.x == 0) fq.idle(L).rewind();
}

// Seed the level-0 frontier with the source (and mark it in the visited
// bitmap, when there is one).
template <class F>
__global__ void k_seed(F q, int32_t* level, int32_t src, Ctl* c, uint32_t* vis) {
  if (threadIdx.x == 0) {
    level[src] = 0;
    if (vis) vis[src >> 5] = 1u << (src & 31);
    c->pending = 1;
    if (!q.push(0, (uint32_t)src)) c->err = 1;
  }
}

// ------------------------------------------------------------------ barrier-free mode
// One frontier queue and no level barrier (GCQ_BFS_ASYNC).  A popped vertex
// u is expanded with the depth it holds now, d(u) + 1 offered to every
// neighbour by an unsigned atomic min on its depth word (-1 reads as the
// largest value, so an unvisited vertex always improves); a neighbour whose
// depth went down is pushed again.  Depths only fall, and every fall is
// followed by a push whose expansion reads the lowered value, so at
// quiescence every edge (u, w) has depth(w) <= depth(u) + 1 and depth(src) =
// 0: the depths are the BFS depths, bit for bit, whatever the order of
// work.  Quiescence: `pending` counts items pushed and not yet fully
// expanded.  A wave adds the items it is about to push before pushing them
// and gives back the items it popped only with its last push of the batch
// (or after the batch when it pushes nothing), so pending > 0 while any item
// is queued or being expanded, and a wave that finds the queue empty may
// stop once it reads pending == 0.  pending also bounds the items in the
// ring, which is checked against the ring's capacity before each push.
constexpr uint32_t kAsyncIdleLimit = 1u << 22;

template <class F>
__device__ __forceinline__ void async_flush(const F& q, uint32_t tid, uint32_t* stage, uint32_t& sc, uint32_t n,
                                            uint32_t& owed, Ctl* c, uint64_t cap) {
  const uint32_t lane = lane_id();
  uint32_t ok = 1;
  if (lane == 0) {
    const uint64_t delta = (uint64_t)n - (uint64_t)owed;  // two's complement: may give items back
    const uint64_t old = gcq::faa_rlx(&c->pending, delta);
    if (n && old + delta > cap) {  // the ring could overflow: give the n back and fail the run
      gcq::faa_rlx(&c->pending, (uint64_t)0 - n);
      gcq::or_rlx(&c->err, 1u);
      ok = 0;
    }
  }
  owed = 0;
  if (!n) return;
  if (gcq::bcast32(ok, 0)) {
    flush(q, tid, stage, sc, n, &c->err);
  } else {
    const uint32_t rest = sc - n;
    const uint32_t mv = lane < rest ? stage[64 + lane] : 0u;
    if (lane < rest) stage[lane] = mv;
    sc = rest;
  }
}

template <class F>
__global__ void k_async(F q, const int32_t* __restrict__ offs, const int32_t* __restrict__ idx, int32_t* level,
                        Ctl* c, uint64_t cap) {
  __shared__ uint32_t stage[16][128];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = lane_id();
  uint32_t* const st = stage[(threadIdx.x / 64) & 15];
  uint32_t* const dep = (uint32_t*)level;
  uint64_t edges = 0, disc = 0;
  uint32_t idle = 0;
  for (;;) {
    uint32_t u = 0;
    const bool got = q.pop(tid, &u);
    const uint64_t gm = __ballot(got);
    if (gm == 0) {
      uint32_t stop = 0;
      if (lane == 0) {
        stop = gcq::ld_rlx(&c->pending) == 0 || gcq::ld_rlx(&c->lbar.abort) != 0;
        if (!stop && ++idle > kAsyncIdleLimit) {
          gcq::st_rlx(&c->lbar.abort, 1u);
          stop = 1;
        }
      }
      if (gcq::bcast32(stop, 0)) break;
      const uint32_t nap = gcq::bcast32(idle, 0);
      for (uint32_t i = 0; i < (nap < 8 ? nap : 8u); ++i) __builtin_amdgcn_s_sleep(2);
      continue;
    }
    idle = 0;
    uint32_t owed = (uint32_t)__builtin_popcountll(gm);  // popped items, given back with the last push
    int32_t beg = 0, deg = 0;
    uint32_t nd = 0;
    if (got) {
      beg = offs[u];
      deg = offs[u + 1] - beg;
      nd = gcq::ld_rlx(&dep[u]) + 1u;  // the depth now: it may have fallen since the push
    }
    edges += (uint64_t)deg;
    int32_t incl = deg;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if ((int)lane >= o) incl += t;
    }
    const int32_t excl = incl - deg;
    const int32_t total = __shfl(incl, 63, 64);
    const bool probe = total > kProbeEdges;
    // this lane's edge target in a round and the depth its owner offers
    auto neighbour = [&](int32_t base, uint32_t* offer) -> uint32_t {
      const int32_t e = base + (int32_t)lane;
      const int32_t ec = e < total ? e : total - 1;
      int lo = 0, hi = 63;
      for (int s2 = 0; s2 < 6; ++s2) {
        const int mid = (lo + hi) >> 1;
        const int32_t pv = __shfl(incl, mid, 64);
        if (pv > ec) hi = mid;
        else lo = mid + 1;
      }
      const int32_t obeg = __shfl(beg, lo, 64);
      const int32_t oexcl = __shfl(excl, lo, 64);
      *offer = (uint32_t)__shfl((int)nd, lo, 64);
      return e < total ? (uint32_t)idx[obeg + (e - oexcl)] : 0u;
    };
    uint32_t sc = 0;
    uint32_t off_next = 0;
    uint32_t w_next = total > 0 ? neighbour(0, &off_next) : 0u;
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t e = base + (int32_t)lane;
      const uint32_t w = w_next, offer = off_next;
      if (base + 64 < total) w_next = neighbour(base + 64, &off_next);
      bool imp = false;
      if (e < total && (!probe || gcq::ld_cached(&dep[w]) > offer)) {
        const uint32_t old = __hip_atomic_fetch_min(&dep[w], offer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        imp = old > offer;
        disc += old == 0xFFFFFFFFu ? 1u : 0u;
      }
      const uint64_t dm = __ballot(imp);
      if (imp) st[sc + gcq::rank_in(dm)] = w;
      sc += (uint32_t)__builtin_popcountll(dm);
      if (sc >= 64) async_flush(q, tid, st, sc, 64, owed, c, cap);
    }
    async_flush(q, tid, st, sc, sc, owed, c, cap);
  }
  edges = gcq::wave_sum(edges);
  disc = gcq::wave_sum(disc);
  if (lane == 0) {
    if (edges) gcq::faa_rlx(&c->edges, edges);
    if (disc) gcq::faa_rlx(&c->reached, disc);
  }
}

// Levels of a finished barrier-free run: the largest depth + 1.
__global__ void k_levels(const int32_t* level, uint64_t n, Ctl* c) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, stp = (uint64_t)gridDim.x * blockDim.x;
  int32_t m = -1;
  for (uint64_t i = gid; i < n; i += stp) m = max(m, level[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if (lane_id() == 0 && m >= 0) __hip_atomic_fetch_max(&c->levels, m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace bfsk

// ======================================================================== host side
namespace {

// The visited bitmap: one bit per vertex, in whole 128-B lines.
uint64_t vis_bytes(uint64_t nv) { return (nv + 1023) / 1024 * 128; }

int ctx_free(gcq_graph* g) {
  BfsCtx& c = g->ctx;
  c.fs.release();
  if (c.level) (void)hipFree(c.level);
  if (c.vis) (void)hipFree(c.vis);
  if (c.ctl) (void)hipFree(c.ctl);
  if (c.hub) (void)hipFree(c.hub);
  if (c.pinned) (void)hipHostFree(c.pinned);
  c = BfsCtx();
  return GCQ_OK;
}

int ctx_ensure(gcq_graph* g, int kind, uint32_t threads) {
  BfsCtx& c = g->ctx;
  const uint32_t nv = (uint32_t)g->host.nv;
  int rc = c.fs.ensure(kind, nv, threads);
  if (rc) return rc;
  if (!c.level) {
    GCQ_HIP(hipMalloc(&c.level, 4ull * nv));
    GCQ_HIP(hipMalloc(&c.vis, vis_bytes(nv)));
    GCQ_HIP(hipMalloc(&c.ctl, sizeof(bfsk::Ctl)));
    GCQ_HIP(hipMalloc(&c.hub, 2 * sizeof(gcqf::HubPool)));
    GCQ_HIP(hipHostMalloc((void**)&c.pinned, 64));
  }
  return GCQ_OK;
}

int reset_run(gcq_graph* g, int32_t src, bool bitmap) {
  BfsCtx& c = g->ctx;
  const uint32_t nv = (uint32_t)g->host.nv;
  hipLaunchKernelGGL(gcqf::kf_fill_i32, dim3(gcqf::grid_of(nv)), dim3(256), 0, g->stream, c.level, (uint64_t)nv, -1);
  if (bitmap)
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(vis_bytes(nv) / 8)), dim3(256), 0, g->stream,
                       (uint64_t*)c.vis, vis_bytes(nv) / 8);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, g->stream, (uint64_t*)c.ctl,
                     (uint64_t)(sizeof(bfsk::Ctl) / 8));
  for (int p = 0; p < 2; ++p)
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, g->stream, (uint64_t*)((gcqf::HubPool*)c.hub + p),
                       gcqf::kHubHeaderWords);
  c.fs.reset(g->stream);
  c.fs.dispatch([&](const auto& fq) {
    using F = std::decay_t<decltype(fq.q[0])>;
    hipLaunchKernelGGL(bfsk::k_seed<F>, dim3(1), dim3(64), 0, g->stream, fq.q[0], c.level, src, (bfsk::Ctl*)c.ctl,
                       bitmap ? c.vis : nullptr);
    return GCQ_OK;
  });
  GCQ_HIP(hipGetLastError());
  return GCQ_OK;
}

template <class F>
int traverse(gcq_graph* g, const bfsk::Frontiers<F>& fq, const gcq_bfs_config* cfg, double* kernel_ms) {
  BfsCtx& c = g->ctx;
  bfsk::Ctl* ctl = (bfsk::Ctl*)c.ctl;
  const uint32_t grid = cfg->threads / cfg->block;
  hipEvent_t e0, e1;
  GCQ_HIP(hipEventCreate(&e0));
  GCQ_HIP(hipEventCreate(&e1));
  int rc = GCQ_OK;
  if (cfg->mode == GCQ_BFS_ASYNC) {
    if constexpr (F::kQueues == 3) {
      gcqh::set_error("the barrier-free BFS needs a ring frontier (G-WFQ, G-LFQ or SFQ); WFQ cells are not reused");
      rc = GCQ_EINVAL;
    } else {
      const uint32_t nv = (uint32_t)g->host.nv;
      GCQ_HIP(hipEventRecord(e0, g->stream));
      hipLaunchKernelGGL(bfsk::k_async<F>, dim3(grid), dim3(cfg->block), 0, g->stream, fq.q[0], g->d_off, g->d_idx,
                         c.level, ctl, (uint64_t)c.fs.capacity);
      GCQ_HIP(hipEventRecord(e1, g->stream));
      hipLaunchKernelGGL(bfsk::k_levels, dim3(gcqf::grid_of(nv)), dim3(256), 0, g->stream, c.level, (uint64_t)nv,
                         ctl);
    }
  } else if (cfg->mode == GCQ_BFS_PERSISTENT || cfg->mode == GCQ_BFS_BITMAP) {
    const bool bitmap = cfg->mode == GCQ_BFS_BITMAP;
    // Long adjacencies are shared out by the grid only where the graph has
    // one: without a pool, a level costs no look at it.
    if (c.max_deg < 0) {
      int32_t m = 0;
      const auto& o = g->host.offsets;
      for (size_t v = 0; v + 1 < o.size(); ++v) m = std::max(m, o[v + 1] - o[v]);
      c.max_deg = m;
    }
    gcqf::HubPool* const hub = c.max_deg > gcqf::kHubDeg ? (gcqf::HubPool*)c.hub : nullptr;
    int per_cu = 0, dev = 0, cus = 0;
    GCQ_HIP(hipGetDevice(&dev));
    GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, bitmap ? bfsk::k_persistent<F, true> : bfsk::k_persistent<F, false>, cfg->block, 0));
    if ((int64_t)grid > (int64_t)per_cu * cus) {
      gcqh::set_error("persistent BFS needs all %u blocks resident; the device holds %d", grid, per_cu * cus);
      rc = GCQ_EINVAL;
    } else {
      // One push per vertex (its claim is won once), into rings of at least
      // V slots reset before the run: the SFQ frontier skips its slot checks.
      auto fqp = fq;
      if constexpr (std::is_same_v<F, gcqf::FSfq>)
        if (c.fs.capacity >= (uint64_t)g->host.nv)
          for (auto& x : fqp.q) x.single_lap = 1u;
      // Every pop of these modes is counted: G-LFQ pushes store first-lap entries.
      if constexpr (std::is_same_v<F, gcqf::FLfq>)
        for (auto& x : fqp.q) x.counted_push = 1u;
      GCQ_HIP(hipEventRecord(e0, g->stream));
      if (bitmap)
        hipLaunchKernelGGL((bfsk::k_persistent<F, true>), dim3(grid), dim3(cfg->block), 0, g->stream, fqp, g->d_off,
                           g->d_idx, c.level, ctl, c.vis, hub);
      else
        hipLaunchKernelGGL((bfsk::k_persistent<F, false>), dim3(grid), dim3(cfg->block), 0, g->stream, fqp,
                           g->d_off, g->d_idx, c.level, ctl, (uint32_t*)nullptr, hub);
      GCQ_HIP(hipEventRecord(e1, g->stream));
    }
  } else {
    // Host level loop (bfs_bench.cpp:583-623): one launch per level, the
    // next frontier size read back after each.
    GCQ_HIP(hipEventRecord(e0, g->stream));
    int32_t L = 0;
    for (;; ++L) {
      hipLaunchKernelGGL(bfsk::k_level<F>, dim3(grid), dim3(cfg->block), 0, g->stream, fq, g->d_off, g->d_idx, c.level,
                         L, ctl);
      if (F::kQueues == 3) hipLaunchKernelGGL(bfsk::k_rewind<F>, dim3(1), dim3(64), 0, g->stream, fq, L);
      GCQ_HIP(hipMemcpyAsync(c.pinned, &ctl->next[L & 1], 4, hipMemcpyDeviceToHost, g->stream));
      GCQ_HIP(hipStreamSynchronize(g->stream));
      if (*(volatile uint32_t*)c.pinned == 0) break;
    }
    GCQ_HIP(hipEventRecord(e1, g->stream));
    c.pinned[1] = (uint32_t)(L + 1);  // levels, as the reference counts them (bfs_bench.cpp:622-631)
    GCQ_HIP(hipMemcpyAsync(&ctl->levels, &c.pinned[1], 4, hipMemcpyHostToDevice, g->stream));
  }
  if (rc == GCQ_OK) {
    GCQ_HIP(hipEventSynchronize(e1));
    GCQ_HIP(hipGetLastError());
    *kernel_ms = gcqh::ms_between(e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int finish_create(gcq_graph* g, gcq_graph** out) {
  *out = g;
  return GCQ_OK;
}

}  // namespace

int gcqg::ensure_uploaded(gcq_graph* g) {
  if (g->d_off) return GCQ_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  const int64_t ne = g->host.ne();
  GCQ_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  GCQ_HIP(hipMalloc(&g->d_off, 4ull * (g->host.nv + 1)));
  GCQ_HIP(hipMalloc(&g->d_idx, 4ull * std::max<int64_t>(ne, 1)));
  GCQ_HIP(hipMemcpy(g->d_off, g->host.offsets.data(), 4ull * (g->host.nv + 1), hipMemcpyHostToDevice));
  if (ne) GCQ_HIP(hipMemcpy(g->d_idx, g->host.indices.data(), 4ull * ne, hipMemcpyHostToDevice));
  return GCQ_OK;
}

extern "C" int gcq_graph_lattice(int32_t rows, int32_t cols, double keep, double diag, uint64_t seed,
                                 gcq_graph** out) {
  GCQ_REQUIRE(out, "null out");
  GCQ_REQUIRE(keep >= 0 && keep <= 1 && diag >= 0 && diag <= 1, "probabilities must be in [0,1]");
  gcq_graph* g = new gcq_graph();
  if (!gcqg::lattice(rows, cols, keep, diag, seed, &g->host)) {
    delete g;
    GCQ_REQUIRE(false, "bad lattice size %dx%d", rows, cols);
  }
  return finish_create(g, out);
}

extern "C" int gcq_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                              gcq_graph** out) {
  GCQ_REQUIRE(out, "null out");
  GCQ_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1, "bad R-MAT probabilities");
  gcq_graph* g = new gcq_graph();
  if (!gcqg::rmat(scale, edge_factor, a, b, c, seed, &g->host)) {
    delete g;
    GCQ_REQUIRE(false, "bad R-MAT scale %d / edge factor %d", scale, edge_factor);
  }
  return finish_create(g, out);
}

extern "C" int gcq_graph_load(const char* path, int32_t format, int32_t undirected, gcq_graph** out) {
  GCQ_REQUIRE(out && path, "null argument");
  GCQ_REQUIRE(format >= GCQ_GRAPH_AUTO && format <= GCQ_GRAPH_SNAP, "format must be GCQ_GRAPH_AUTO, _MTX or _SNAP");
  gcq_graph* g = new gcq_graph();
  std::string err;
  if (gcqg::load_edge_file(path, format, undirected != 0, &g->host, &err) != 0) {
    delete g;
    GCQ_REQUIRE(false, "%s", err.c_str());
  }
  return finish_create(g, out);
}

extern "C" int gcq_graph_from_csr(int32_t nv, int64_t ne, const int32_t* offsets, const int32_t* indices,
                                  gcq_graph** out) {
  GCQ_REQUIRE(out && offsets && (indices || ne == 0) && nv > 0 && ne >= 0 && ne < (1ll << 31), "bad CSR");
  GCQ_REQUIRE(offsets[0] == 0 && offsets[nv] == ne, "offsets must start at 0 and end at ne");
  for (int32_t v = 0; v < nv; ++v)
    GCQ_REQUIRE(offsets[v] <= offsets[v + 1], "offsets decrease at vertex %d", v);
  gcq_graph* g = new gcq_graph();
  g->host.nv = nv;
  g->host.offsets.assign(offsets, offsets + nv + 1);
  g->host.indices.assign(indices, indices + ne);
  for (int64_t i = 0; i < ne; ++i)
    if (indices[i] < 0 || indices[i] >= nv) {
      delete g;
      GCQ_REQUIRE(false, "index %lld out of range", (long long)i);
    }
  return finish_create(g, out);
}

extern "C" int gcq_graph_info(const gcq_graph* g, int32_t* nv, int64_t* ne) {
  GCQ_REQUIRE(g, "null graph");
  if (nv) *nv = g->host.nv;
  if (ne) *ne = g->host.ne();
  return GCQ_OK;
}

extern "C" int gcq_graph_copy_csr(const gcq_graph* g, int32_t* offsets, int32_t* indices) {
  GCQ_REQUIRE(g, "null graph");
  if (offsets) std::memcpy(offsets, g->host.offsets.data(), 4ull * (g->host.nv + 1));
  if (indices) std::memcpy(indices, g->host.indices.data(), 4ull * g->host.ne());
  return GCQ_OK;
}

extern "C" int32_t gcq_graph_pick_source(const gcq_graph* g, uint64_t seed, uint64_t k) {
  return g ? gcqg::pick_source(g->host, seed, k) : -1;
}

extern "C" int gcq_graph_destroy(gcq_graph* g) {
  if (!g) return GCQ_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  ctx_free(g);
  if (g->d_off) (void)hipFree(g->d_off);
  if (g->d_idx) (void)hipFree(g->d_idx);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return GCQ_OK;
}

extern "C" int gcq_bfs_run(gcq_graph* g, int32_t src, const gcq_bfs_config* cfg, int32_t* depth_out,
                           gcq_bfs_result* res) {
  GCQ_REQUIRE(g && cfg && res, "null argument");
  GCQ_REQUIRE(cfg->block >= 64 && cfg->block % 64 == 0 && cfg->block <= 1024, "block must be a multiple of 64");
  GCQ_REQUIRE(cfg->threads >= cfg->block && cfg->threads % cfg->block == 0,
              "threads must be a positive multiple of block (whole waves, bfs_bench.cpp:521)");
  GCQ_REQUIRE(cfg->mode == GCQ_BFS_HOST_LOOP || cfg->mode == GCQ_BFS_PERSISTENT || cfg->mode == GCQ_BFS_ASYNC ||
                  cfg->mode == GCQ_BFS_BITMAP,
              "bad mode");
  std::memset(res, 0, sizeof *res);
  if (src < 0 || src >= g->host.nv) src = 0;  // bfs_bench.cpp:555
  int rc = gcqg::ensure_uploaded(g);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  rc = ctx_ensure(g, cfg->kind, cfg->threads);
  if (rc) return rc;
  rc = reset_run(g, src, cfg->mode == GCQ_BFS_BITMAP);
  if (rc) return rc;
  BfsCtx& c = g->ctx;
  double kms = 0;
  rc = c.fs.dispatch([&](const auto& fq) { return traverse(g, fq, cfg, &kms); });
  if (rc) return rc;
  bfsk::Ctl h;
  GCQ_HIP(hipMemcpyAsync(&h, c.ctl, sizeof h, hipMemcpyDeviceToHost, g->stream));
  if (depth_out)
    GCQ_HIP(hipMemcpyAsync(depth_out, c.level, 4ull * g->host.nv, hipMemcpyDeviceToHost, g->stream));
  GCQ_HIP(hipStreamSynchronize(g->stream));
  res->edges_scanned = h.edges;
  res->reached = h.reached + 1;
  res->levels = h.levels;
  res->kernel_ms = kms;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (h.lbar.abort) {
    gcqh::set_error("BFS aborted: a grid-barrier wait exceeded its spin limit");
    return GCQ_ETIMEOUT;
  }
  if (h.err) {
    gcqh::set_error("BFS frontier enqueue failed (queue capacity or cell pool exhausted)");
    return GCQ_EFULL;
  }
  return GCQ_OK;
}

// ======================================================================== allocation-free BFS
// A BFS over a caller's device-resident CSR into a caller's device depth
// array that allocates nothing: the two G-LFQ frontier rings and the control
// block live in `scratch`, a file-scope device array each GPU holds until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name stands for a host-side shadow, not a device address); it
// uses the per-device address hipGetSymbolAddress returned.  The kernels are
// the ones gcq_bfs_run launches for G-LFQ frontiers in persistent mode.
namespace {

constexpr int32_t kNoAllocMaxVertices = 1 << 21;
constexpr uint64_t kNoAllocRingBytes = gcq::lfq::ring_bytes((uint32_t)kNoAllocMaxVertices);
constexpr uint64_t kNoAllocCtlOff = 2 * kNoAllocRingBytes;
constexpr uint64_t kNoAllocBytes = kNoAllocCtlOff + ((sizeof(bfsk::Ctl) + 255) & ~255ull);
static_assert(kNoAllocRingBytes % 256 == 0, "ring alignment");
static_assert(kNoAllocBytes == GCQ_BFS_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array, looked up once per
// device (with that device current).
int scratch_for_current_device(void** out) {
  int dev = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d devices supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GCQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GCQ_REQUIRE(p != nullptr, "no device address for the BFS scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GCQ_OK;
}

}  // namespace

extern "C" int gcq_bfs_run_noalloc(const int32_t* d_offsets, const int32_t* d_indices, int32_t nv, int64_t ne,
                                   int32_t src, uint32_t threads, uint32_t block, void* stream, int32_t* d_depth,
                                   gcq_bfs_result* res) {
  GCQ_REQUIRE(d_offsets && d_depth && res && (d_indices || ne == 0), "null argument");
  GCQ_REQUIRE(nv >= 1 && nv <= kNoAllocMaxVertices,
              "nv = %d: the allocation-free BFS takes 1 to %d vertices (its frontier rings are fixed)", nv,
              kNoAllocMaxVertices);
  GCQ_REQUIRE(ne >= 0 && ne < (1ll << 31), "ne out of range");
  GCQ_REQUIRE(block >= 64 && block % 64 == 0 && block <= 1024, "block must be a multiple of 64");
  GCQ_REQUIRE(threads >= block && threads % block == 0 && threads <= (uint32_t)kNoAllocMaxVertices,
              "threads must be a positive multiple of block, at most %d", kNoAllocMaxVertices);
  std::memset(res, 0, sizeof *res);
  if (src < 0 || src >= nv) src = 0;  // bfs_bench.cpp:555
  void* base = nullptr;
  int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  int dev = 0, per_cu = 0, cus = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bfsk::k_persistent<gcqf::FLfq>, block, 0));
  const uint32_t grid = threads / block;
  GCQ_REQUIRE((int64_t)grid <= (int64_t)per_cu * cus, "persistent BFS needs all %u blocks resident; the device holds %d",
              grid, per_cu * cus);
  hipStream_t s = (hipStream_t)stream;
  const uint32_t cap = gcqf::pow2_at_least(std::max<uint64_t>((uint64_t)nv, threads));
  void* ring0 = base;
  void* ring1 = (char*)base + kNoAllocRingBytes;
  bfsk::Ctl* ctl = (bfsk::Ctl*)((char*)base + kNoAllocCtlOff);
  bfsk::Frontiers<gcqf::FLfq> fq;
  for (int i = 0; i < 3; ++i) fq.q[i] = gcqf::FLfq{(gcq::lfq::Ring*)(i % 2 ? ring1 : ring0)};
  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(gcqf::kf_fill_i32, dim3(gcqf::grid_of(nv)), dim3(256), 0, s, d_depth, (uint64_t)nv, -1);
  hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(1), dim3(256), 0, s, (uint64_t*)ctl, (uint64_t)(sizeof(bfsk::Ctl) / 8));
  hipLaunchKernelGGL(gcqf::kf_reset_lfq, dim3(gcqf::grid_of(2ull * cap)), dim3(256), 0, s, ring0, ring1, cap);
  hipLaunchKernelGGL(bfsk::k_seed<gcqf::FLfq>, dim3(1), dim3(64), 0, s, fq.q[0], d_depth, src, ctl, (uint32_t*)nullptr);
  hipLaunchKernelGGL(bfsk::k_persistent<gcqf::FLfq>, dim3(grid), dim3(block), 0, s, fq, d_offsets, d_indices, d_depth,
                     ctl, (uint32_t*)nullptr, (gcqf::HubPool*)nullptr);
  GCQ_HIP(hipGetLastError());
  bfsk::Ctl h;
  GCQ_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, s));
  GCQ_HIP(hipStreamSynchronize(s));
  res->edges_scanned = h.edges;
  res->reached = h.reached + 1;
  res->levels = h.levels;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  res->kernel_ms = res->total_ms;  // no events: they would be allocations
  if (h.lbar.abort) {
    gcqh::set_error("BFS aborted: a grid-barrier wait exceeded its spin limit");
    return GCQ_ETIMEOUT;
  }
  if (h.err) {
    gcqh::set_error("BFS frontier enqueue failed");
    return GCQ_EFULL;
  }
  return GCQ_OK;
}
