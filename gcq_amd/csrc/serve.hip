// serve.hip — continuously batched serving of BFS queries and ray-traced
// tile jobs in one persistent kernel.
//
// No counterpart in the reference (its harnesses run one workload per
// process).  Per request the results are those of the plain paths: a BFS
// query's depths are exact BFS depths (bfs_bench.cpp:327-633 semantics,
// shared expansion in bfs_batch.hpp) and a tile's pixels follow
// rt_final.cpp's image (rt_shade.hpp), so each result is the same alone and
// in the mix.
//
// Design (MI355X):
//  * One launch runs the whole service.  Its first bfs_blocks blocks form a
//    BFS sub-grid: every admitted query advances one level per grid barrier
//    through a shared (slot, vertex) frontier queue, as in batch.hip.  The
//    remaining blocks form an RT sub-grid: each wave picks an active tile
//    job, claims 64 primary rays at a time or pops a live bounce ray from
//    the job's queue, and traces one bounce.
//  * Requests reach the kernel through two device-resident submission
//    queues in HBM, one for BFS queries and one for tile jobs: G-LFQ rings
//    of slot numbers beside a per-slot request descriptor.  The host's only
//    channel in is a doorbell ring in pinned memory (an entry per request,
//    its sequence word written last).  One gate wave, the first wave after
//    the BFS sub-grid, reads new doorbell entries (system-scope loads), copies
//    each request into its slot's descriptor and pushes the slot onto its
//    queue; nothing else in the kernel reads host memory on the way in.
//  * The persistent consumers pull from those queues: the last block to
//    reach each BFS barrier pops BFS queries and installs them into a slot
//    whose depth array the sub-grid clears during the next level and whose
//    source enters the frontier for the level after; an RT wave that finds
//    no work, and every RT wave once in kRtSubmitEvery passes, pops tile
//    jobs and installs them into an RT job slot at once, so a tile never
//    waits for the BFS level cadence.  Completions go back through a pinned
//    completion ring.
//  * The host admission loop decides when a request may enter: a free slot,
//    room in the device-memory budget (depth arrays, ray records, images)
//    and in the frontier ring (the sum of V of in-flight queries).
//  * Nothing is launched while the server runs: the host only reads and
//    writes pinned memory, so no stream can end up queued behind the
//    persistent kernel on a shared hardware queue.
#pragma clang fp contract(off)
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <type_traits>
#include <vector>

#include "bfs_batch.hpp"
#include "graph_handle.hpp"
#include "rt_handle.hpp"

namespace sv {

using namespace gcqf;

constexpr uint32_t kMaxAdmit = 128;          // BFS queries installed per BFS barrier (a burst joins its groups in fewer levels: 300-request burst +9% against 32)
// An idle BFS lane sleeps 32 x 127 x 64 cycles (about 0.1 ms) per level.
// Measured on the served mix (Poisson, 20 and 150 requests/s) against no
// park: small-graph p50 2.3 -> 0.9 ms and tile p50 0.58 -> 0.18 ms at
// 20/s, overall p50 3.3 -> 2.0 ms at 150/s; burst capacity unchanged.
// Shorter parks (4, 16) kept more of the interference, 64 gained nothing.
constexpr uint32_t kIdleParkSleeps = 32;
constexpr uint64_t kRtIdleLimit = 1ull << 30;  // idle polls of an RT wave before it gives up
constexpr uint32_t kMaxTilePixels = 1u << 16;  // RT items carry a 16-bit pixel and a 15-bit job generation
constexpr uint32_t kRtSubmitEvery = 32;        // passes of a busy RT wave between looks at the tile queue

// Doorbell entry (pinned host memory, written by the host; seq last).  The
// gate wave copies it into its slot's descriptor in HBM (the same layout).
struct alignas(64) Req {
  uint64_t seq;  // i + 1 once entry i is complete
  uint64_t id;   // server-internal request number
  int32_t type, graph, src, slot;
  int32_t x0, y0, w, h;
  uint64_t buf, buf2;  // BFS: depth array, visited bitmap; RT: image, ray records (device arena)
};
static_assert(sizeof(Req) == 64, "req");

// Completion ring entry (pinned host memory, written by the kernel; seq last).
struct alignas(64) Comp {
  uint64_t seq;
  uint64_t id;
  int32_t type, status, levels, slot;
  uint64_t rays;
  uint64_t pad[3];
};
static_assert(sizeof(Comp) == 64, "comp");

struct GraphDev {
  uint64_t off, idx;
  uint32_t nv, lane;  // lane: the BFS lane that serves this graph's queries
  uint64_t pad2;
};

// One RT job slot; the three counters sit on their own 128-B lines.
struct RtJob {
  uint64_t img, recs, id;
  int32_t x0, y0, w, h;
  uint32_t gen, npx, active, pad;
  uint64_t p0[9];
  uint64_t next_primary, p1[15];  // gen << 32 | next unclaimed primary ray
  uint64_t alive, p2[15];         // live rays
  uint64_t traced, p3[15];
};
static_assert(sizeof(RtJob) == 512, "rt job");

// One BFS lane: a sub-grid with its own level barrier, frontier queues and
// submission queue.  The second lane, when configured, takes the queries on
// graphs of at most 2^small_log2 vertices, so that they advance at their
// own level cadence instead of waiting out the big graphs' levels.
constexpr int kLanes = 2;
struct LaneCtl {
  Barrier bar;
  uint32_t adm_n, adm_slot[kMaxAdmit], p3[31];
  int32_t active, p6[31];
  int32_t levels, p8[31];
  // Set only by the lane's own last block inside its barrier, so every block
  // of the lane reads the same value after that barrier (the shared exit
  // word can be set by the other lane at any time).
  uint32_t quit, p9[31];
};

struct Ctl {
  LaneCtl lane[kLanes];
  uint32_t err, p1[31];  // 1 frontier enqueue, 2 RT enqueue, 4 completion ring full, 8 RT idle limit
  uint32_t exit, p2[31];
  uint64_t sub_head, p4[15];  // doorbell entries the gate wave has moved
  uint32_t stop_seen, p9[31];  // the gate wave saw the host's stop word
  uint64_t comp_tail, p5[15];
  int32_t active_rt, p7[31];
};

// Multi-source groups (gcq_server_config.multi_source): per group, its
// members and the members settled at a level of each parity, on one line;
// written at admission and by done_scan (one thread per group), gathered
// into by the expansion.
struct GroupCtl {
  uint64_t members;   // busy bits
  uint64_t fresh[2];  // bk::Group.fresh points here
  uint32_t lane, nv;
  uint64_t pad[12];
};
static_assert(sizeof(GroupCtl) == 128, "group ctl");

struct Dev {
  bk::Dev bd;              // read side of the BFS slot table
  bk::Slot* slots;         // the same table, written at admission
  uint32_t* slot_active;   // per BFS slot: 1 + the lane serving it, 0 when free
  const GraphDev* graphs;
  RtJob* jobs;
  char* rings;             // one G-LFQ ring per RT job slot
  uint64_t ring_stride;
  Req* sub;                // doorbell ring (pinned host memory)
  gcq::lfq::Ring* subq[3]; // device-resident submission queues: BFS slots (lane 0, lane 1), RT job slots
  Req* desc[2];            // per-slot request descriptors in HBM: BFS, RT
  Comp* comp;
  const uint32_t* stop;    // set by the host once every request has completed
  Ctl* c;
  rtk::Scene sc;
  int32_t W, H, max_depth;
  uint32_t njobs, bfs_blocks, max_requests, rt_workers;  // rt_workers: the RT sub-grid runs tile jobs
  uint32_t lane_blocks[kLanes];  // blocks [0, lane_blocks[0]) are lane 0, the next lane_blocks[1] lane 1
  // multi-source groups (ms != 0): the group table, its control lines and
  // per member (64 per group) the depth array, start level and BFS slot
  uint32_t ms, ngroups;
  bk::MsDev md;
  bk::Group* groups;
  GroupCtl* gctl;
  uint64_t* gdepth;
  int32_t* gstart;
  uint32_t* gslot;
  bk::HubPool* hub;  // per BFS lane, the two level-parity pools of long adjacencies (multi-source form)
};

__device__ void publish(const Dev& d, uint64_t id, int32_t type, int32_t levels, int32_t slot, uint64_t rays) {
  const uint64_t k = gcq::faa_rlx(&d.c->comp_tail, (uint64_t)1);
  if (k >= d.max_requests) {
    gcq::or_rlx(&d.c->err, 4u);
    return;
  }
  Comp* e = &d.comp[k];
  gcq::st_sys(&e->id, id);
  gcq::st_sys(&e->type, type);
  gcq::st_sys(&e->status, 0);
  gcq::st_sys(&e->levels, levels);
  gcq::st_sys(&e->slot, slot);
  gcq::st_sys(&e->rays, rays);
  gcq::vm_wait();  // the record before its sequence word
  gcq::st_sys(&e->seq, k + 1);
}

// The gate wave: move each complete doorbell entry into its slot's
// descriptor and push the slot onto the BFS or RT submission queue, until
// the kernel exits.  Lane 0 does the work.  The host writes an entry only
// for a slot it has freed, so a slot is in at most one queue at a time and
// each queue holds at most its slot count (its capacity).
__device__ void gate(const Dev& d) {
  Ctl* c = d.c;
  uint64_t h = 0;
  for (;;) {
    uint32_t moved = 0, quit = 0;
    if (gcq::lane_id() == 0) {
      while (h < d.max_requests) {
        const Req* e = &d.sub[h];
        if (gcq::ld_sys(&e->seq) != h + 1) break;
        gcq::vm_wait();  // the sequence word before the fields
        const int32_t type = gcq::ld_sys(&e->type), slot = gcq::ld_sys(&e->slot), graph = gcq::ld_sys(&e->graph);
        const int t = type == GCQ_REQ_BFS ? 0 : 1;
        const int q = t == 0 ? (int)d.graphs[graph].lane : kLanes;
        Req* dst = &d.desc[t][slot];
        gcq::st_rlx(&dst->id, gcq::ld_sys(&e->id));
        gcq::st_rlx(&dst->type, type);
        gcq::st_rlx(&dst->graph, graph);
        gcq::st_rlx(&dst->src, gcq::ld_sys(&e->src));
        gcq::st_rlx(&dst->slot, slot);
        gcq::st_rlx(&dst->x0, gcq::ld_sys(&e->x0));
        gcq::st_rlx(&dst->y0, gcq::ld_sys(&e->y0));
        gcq::st_rlx(&dst->w, gcq::ld_sys(&e->w));
        gcq::st_rlx(&dst->h, gcq::ld_sys(&e->h));
        gcq::st_rlx(&dst->buf, gcq::ld_sys(&e->buf));
        gcq::st_rlx(&dst->buf2, gcq::ld_sys(&e->buf2));
        // enqueue waits for these stores before it publishes the slot
        gcq::lfq::enqueue(d.subq[q], (uint32_t)slot);
        ++h;
        ++moved;
      }
      if (moved) gcq::st_rlx(&c->sub_head, h);
      else if (gcq::ld_sys(d.stop)) gcq::st_rlx(&c->stop_seen, 1u);
      quit = gcq::ld_rlx(&c->exit) | gcq::ld_rlx(&c->lane[0].bar.abort) | gcq::ld_rlx(&c->lane[1].bar.abort);
    }
    if (gcq::bcast32(quit, 0)) break;
    if (!gcq::bcast32(moved, 0)) __builtin_amdgcn_s_sleep(8);
  }
}

// Install a lane's queued BFS queries (one thread: the last block to reach
// the barrier that ends the lane's level L).  True when the lane is idle:
// nothing admitted here and no query of the lane still running, so its next
// level has no items.
__device__ bool admit_bfs(const Dev& d, int lane, int32_t L) {
  Ctl* c = d.c;
  LaneCtl* lc = &c->lane[lane];
  uint32_t n = 0;
  while (n < kMaxAdmit) {
    const uint32_t slot = gcq::lfq::dequeue(d.subq[lane], (gcq::lfq::NoProbe*)nullptr, true);
    if (slot == gcq::lfq::kBot) break;
    const Req* e = &d.desc[0][slot];
    const GraphDev& g = d.graphs[gcq::ld_rlx(&e->graph)];
    bk::Slot* s = &d.slots[slot];
    gcq::st_rlx(&s->off, g.off);
    gcq::st_rlx(&s->idx, g.idx);
    gcq::st_rlx(&s->depth, gcq::ld_rlx(&e->buf));
    gcq::st_rlx(&s->visited, gcq::ld_rlx(&e->buf2));
    gcq::st_rlx(&s->src, (uint32_t)gcq::ld_rlx(&e->src));
    gcq::st_rlx(&s->nv, g.nv);
    gcq::st_rlx(&s->start, (uint32_t)(L + 2));  // cleared during L+1, source expanded at L+2
    gcq::st_rlx(&s->pad2[0], gcq::ld_rlx(&e->id));
    if (d.ms) {  // the host chose the member: group in x0, bit in y0
      const uint32_t gi = (uint32_t)gcq::ld_rlx(&e->x0), b = (uint32_t)gcq::ld_rlx(&e->y0);
      gcq::st_rlx(&s->pad2[1], ((uint64_t)gi << 8) | b);
      gcq::st_rlx(&d.gdepth[64ull * gi + b], gcq::ld_rlx(&e->buf));
      gcq::st_rlx(&d.gstart[64ull * gi + b], L + 2);
      gcq::st_rlx(&d.gslot[64ull * gi + b], slot);
      gcq::st_rlx(&d.gctl[gi].members, gcq::ld_rlx(&d.gctl[gi].members) | (1ull << b));
    }
    gcq::st_rlx(&d.bd.alive[slot * bk::kAliveStride], (uint32_t)(L + 2));
    gcq::st_rlx(&d.slot_active[slot], 1u + (uint32_t)lane);
    gcq::st_rlx(&lc->adm_slot[n], slot);
    ++n;
    gcq::faa_rlx(&lc->active, 1);
  }
  gcq::st_rlx(&lc->adm_n, n);
  // The host sets its stop word only once every request has completed, and
  // the gate wave reports it only after finding no doorbell entry left.
  const bool idle = n == 0 && gcq::ld_rlx(&c->stop_seen) && gcq::ld_rlx(&c->lane[0].active) == 0 &&
                    gcq::ld_rlx(&c->lane[1].active) == 0 && gcq::ld_rlx(&c->active_rt) == 0;
  // a lane whose barrier gave up (spin limit) takes the whole server down
  const bool aborted = gcq::ld_rlx(&c->lane[0].bar.abort) | gcq::ld_rlx(&c->lane[1].bar.abort);
  if (idle || aborted) {
    gcq::st_rlx(&lc->quit, 1u);
    gcq::st_rlx(&c->exit, 1u);  // for the gate and the RT waves
  }
  return n == 0 && gcq::ld_rlx(&lc->active) == 0;
}

// Install one queued tile job, if any (lane 0 of an RT wave); true if one was.  The job's
// fields land before the claim counter that carries its generation, and
// that before the slot reads as active.
__device__ bool admit_rt(const Dev& d) {
  const uint32_t slot = gcq::lfq::dequeue(d.subq[kLanes], (gcq::lfq::NoProbe*)nullptr, true);
  if (slot == gcq::lfq::kBot) return false;
  const Req* e = &d.desc[1][slot];
  RtJob* j = &d.jobs[slot];
  const uint32_t gen = (gcq::ld_rlx(&j->gen) + 1u) & 0x7FFFu;
  const int32_t w = gcq::ld_rlx(&e->w), hh = gcq::ld_rlx(&e->h);
  gcq::st_rlx(&j->img, gcq::ld_rlx(&e->buf));
  gcq::st_rlx(&j->recs, gcq::ld_rlx(&e->buf2));
  gcq::st_rlx(&j->id, gcq::ld_rlx(&e->id));
  gcq::st_rlx(&j->x0, gcq::ld_rlx(&e->x0));
  gcq::st_rlx(&j->y0, gcq::ld_rlx(&e->y0));
  gcq::st_rlx(&j->w, w);
  gcq::st_rlx(&j->h, hh);
  gcq::st_rlx(&j->npx, (uint32_t)(w * hh));
  gcq::st_rlx(&j->gen, gen);
  gcq::st_rlx(&j->alive, (uint64_t)(w * hh));
  gcq::st_rlx(&j->traced, (uint64_t)0);
  gcq::vm_wait();
  gcq::st_rlx(&j->next_primary, (uint64_t)gen << 32);
  gcq::vm_wait();
  gcq::faa_rlx(&d.c->active_rt, 1);  // counted before the job can complete
  gcq::st_rlx(&j->active, 1u);
  return true;
}

// The lane's first block after the barrier ending its level L: the lane's
// queries with nothing discovered at level L are finished.
__device__ void done_scan(const Dev& d, int lane, int32_t L) {
  for (uint32_t s = threadIdx.x; s < d.bd.nslots; s += blockDim.x) {
    if (gcq::ld_rlx(&d.slot_active[s]) != 1u + (uint32_t)lane) continue;
    const int32_t start = (int32_t)gcq::ld_rlx(&d.slots[s].start);
    if (start > L || (int32_t)gcq::ld_rlx(&d.bd.alive[s * bk::kAliveStride]) >= L + 1) continue;
    const uint64_t id = gcq::ld_rlx(&d.slots[s].pad2[0]);
    gcq::st_rlx(&d.slot_active[s], 0u);
    gcq::faa_rlx(&d.c->lane[lane].active, -1);
    gcq::vm_wait();  // slot free and count down before the host can reuse the slot
    publish(d, id, GCQ_REQ_BFS, L + 1 - start, (int32_t)s, 0);
  }
}

// One BFS lane's sub-grid: blocks [first, first + lane_blocks[lane]).
// Multi-source form of done_scan (the lane's first block after the barrier
// ending its level L, one thread per group): a member with no vertex settled
// at level L is finished; its levels are L - start.  The level's gather word
// is cleared for level L + 2.
__device__ void done_scan_ms(const Dev& d, int lane, int32_t L) {
  for (uint32_t gi = threadIdx.x; gi < d.ngroups; gi += blockDim.x) {
    GroupCtl* gc = &d.gctl[gi];
    if (gcq::ld_rlx(&gc->lane) != (uint32_t)lane) continue;
    const uint64_t act = gcq::ld_rlx(&gc->fresh[L & 1]);
    gcq::st_rlx(&gc->fresh[L & 1], 0ull);
    uint64_t mem = gcq::ld_rlx(&gc->members);
    for (uint64_t bits = mem; bits; bits &= bits - 1) {
      const int b = __builtin_ctzll(bits);
      const int32_t start = gcq::ld_rlx(&d.gstart[64ull * gi + b]);
      if (start > L || ((act >> b) & 1ull)) continue;
      const uint32_t slot = gcq::ld_rlx(&d.gslot[64ull * gi + b]);
      const uint64_t id = gcq::ld_rlx(&d.slots[slot].pad2[0]);
      mem &= ~(1ull << b);
      gcq::st_rlx(&gc->members, mem);
      gcq::st_rlx(&d.slot_active[slot], 0u);
      gcq::faa_rlx(&d.c->lane[lane].active, -1);
      gcq::vm_wait();  // member and slot free, count down, before the host can reuse them
      publish(d, id, GCQ_REQ_BFS, L - start, (int32_t)slot, 0);
    }
  }
}

template <class F>
__device__ void bfs_part(const Frontiers<F>& fq, const Dev& d, int lane, uint32_t first, uint32_t* stage) {
  LaneCtl* const lc = &d.c->lane[lane];
  const uint32_t nblocks = d.lane_blocks[lane];
  const uint32_t tid = (blockIdx.x - first) * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)nblocks * blockDim.x;
  uint64_t prev_end[2] = {kNoHead, kNoHead};  // per frontier: its end when last drained
  int32_t L = 0;
  for (;; ++L) {
    if (F::kQueues == 3) fq.idle(L).recycle(tid, nthreads, tid);
    // Head of the frontier drained at the last barrier back to its end (it
    // is this level's output; counted pops overshoot Head).
    if constexpr (F::kQueues == 2)
      if (L > 0 && blockIdx.x == first && threadIdx.x == 0) fq.in(L - 1).drained(prev_end[(L - 1) & 1]);
    // Queries installed at the previous barrier: clear their depth arrays
    // (0 at the source) and visited bitmaps, and put their sources into this
    // level's output.  Their first expansion is a level later, after the
    // barrier's cache drop below.
    const uint32_t na = gcq::ld_rlx(&lc->adm_n);
    for (uint32_t a = 0; a < na; ++a) {
      const uint32_t s = gcq::ld_rlx(&lc->adm_slot[a]);
      const bk::Slot* sl = &d.slots[s];
      int32_t* dep = (int32_t*)gcq::ld_rlx(&sl->depth);
      const uint32_t nv = gcq::ld_rlx(&sl->nv), src = gcq::ld_rlx(&sl->src);
      if (d.ms) {
        // Clear the member's bit from its group's seen words (the bit's last
        // user may have left it set; other members settle concurrently, so
        // by atomic and), and gather the source into this level's word: it
        // settles at level L + 1 with depth 0.
        const uint64_t gb = gcq::ld_rlx(&sl->pad2[1]);
        const uint32_t gi = (uint32_t)(gb >> 8);
        const uint64_t bit = 1ull << (gb & 63u);
        const bk::Group& G = d.groups[gi];
        uint64_t* seen = (uint64_t*)G.seen;
        for (uint64_t v = tid; v < nv; v += nthreads) {
          gcq::st_rlx(&dep[v], -1);
          if (gcq::ld_rlx(&seen[v]) & bit) __hip_atomic_fetch_and(&seen[v], ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tid == 0 && gcq::or_rlx((uint64_t*)G.nxt[L & 1] + src, bit) == 0ull &&
            !fq.out(L).push(tid, (gi << d.bd.vbits) | src))
          gcq::or_rlx(&d.c->err, 1u);
      } else {
        uint32_t* vis = (uint32_t*)gcq::ld_rlx(&sl->visited);
        for (uint64_t v = tid; v < nv; v += nthreads) gcq::st_rlx(&dep[v], v == src ? 0 : -1);
        if (vis)  // 0 = no bitmap (the slot table's convention; the host always gives one)
          for (uint64_t w = tid; w < ((uint64_t)nv + 31) / 32; w += nthreads)
            gcq::st_rlx(&vis[w], w == src / 32 ? 1u << (src & 31u) : 0u);
        if (tid == 0 && !fq.out(L).push(tid, (s << d.bd.vbits) | src)) gcq::or_rlx(&d.c->err, 1u);
      }
    }
    // Slot fields, depths and bitmaps are read through the caches: they only
    // change at a barrier (admission) or, for a query not yet expanding, in
    // the level after it, and every block drops its CU's and XCD's cached
    // lines after each barrier.
    uint64_t end = 0;
    bk::HubPool* const hp = &d.hub[2 * lane + (L & 1)];
    if (d.ms)
      (void)bk::ms_expand<F, true>(fq.in(L), fq.out(L), d.md, L, tid, stage,
                                   F::kHeadFromDrain ? prev_end[L & 1] : kNoHead, nthreads / 64, &end, hp);
    else
      (void)bk::expand<F, false>(fq.in(L), fq.out(L), d.bd, L, tid, stage,
                                 F::kHeadFromDrain ? prev_end[L & 1] : kNoHead, nthreads / 64, &end, hp);
    prev_end[L & 1] = end;
    auto level_end = [&]() -> bool {
      if (F::kQueues == 3) fq.idle(L).rewind();
      bk::hub_reset(&d.hub[2 * lane + ((L + 1) & 1)]);
      return admit_bfs(d, lane, L);
    };
    // A level that posted long adjacencies ends after the lane's scan of
    // them, at a second barrier.  flag: 1 long adjacencies posted, 2 the
    // lane is idle.
    uint32_t flag = 0;
    bool ok = grid_barrier(&lc->bar, nblocks, [&]() -> uint32_t {
      if (gcq::ld_rlx(&hp->word) != 0ull) return 1u;
      return level_end() ? 2u : 0u;
    }, &flag);
    if (ok) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (flag == 1u) {
        if (d.ms) (void)bk::ms_hubs(fq.out(L), d.md, hp, L, tid, (uint32_t)(nthreads / 64), stage);
        else (void)bk::hubs<F, false>(fq.out(L), d.bd, hp, L, tid, (uint32_t)(nthreads / 64), stage);
        ok = grid_barrier(&lc->bar, nblocks, [&]() { (void)level_end(); });
      }
    }
    if (!ok) {
      if (threadIdx.x == 0) gcq::st_rlx(&d.c->exit, 1u);  // the gate and the RT waves leave too
      break;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (gcq::ld_rlx(&lc->quit)) break;
    if (blockIdx.x == first) {
      if (d.ms) done_scan_ms(d, lane, L);
      else done_scan(d, lane, L);
    }
    // An idle lane parks before its next (empty) level, so that its blocks'
    // barrier atomics, release polls and cache invalidations stay off the
    // memory system while the other lane and the tile waves run; a query
    // then waits at most one park for admission.
    if (flag == 2u)
      for (uint32_t i = 0; i < kIdleParkSleeps; ++i) __builtin_amdgcn_s_sleep(127);
  }
  if (tid == 0) lc->levels = L + 1;
}

__device__ void rt_part(const Dev& d) {
  const uint32_t lane = gcq::lane_id();
  const uint32_t rtid = (blockIdx.x - d.bfs_blocks) * blockDim.x + threadIdx.x;
  const rtk::V3 light = rtk::norm(d.sc.light_raw);
  uint32_t j = (rtid / 64) % d.njobs;
  uint64_t idle = 0;
  uint32_t pass = rtid / 64;  // staggers the busy waves' looks at the tile queue
  for (;;) {
    int32_t pick = -1;
    for (uint32_t k = 0; k < d.njobs; ++k) {
      const uint32_t jj = (j + k) % d.njobs;
      if (gcq::ld_rlx(&d.jobs[jj].active)) {
        pick = (int32_t)jj;
        break;
      }
    }
    if (pick < 0 || ++pass % kRtSubmitEvery == 0) {
      uint32_t got = 0;
      if (lane == 0) got = admit_rt(d) ? 1u : 0u;
      if (gcq::bcast32(got, 0) && pick < 0) continue;  // pick the job just installed
    }
    if (pick < 0) {
      if (gcq::ld_rlx(&d.c->exit)) break;
      if (++idle > kRtIdleLimit) {
        if (lane == 0) gcq::or_rlx(&d.c->err, 8u);
        break;
      }
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    idle = 0;
    j = (uint32_t)pick;
    RtJob* job = &d.jobs[j];
    const FLfq q{(gcq::lfq::Ring*)(d.rings + j * d.ring_stride)};
    // Claim up to 64 primaries; the counter carries the job generation, so
    // a claim made after the slot was reused reloads the new job's fields.
    uint64_t claim = ~0ull;
    if (lane == 0) {
      const uint64_t cur = gcq::ld_rlx(&job->next_primary);
      if ((uint32_t)cur < gcq::ld_rlx(&job->npx)) claim = gcq::faa_rlx(&job->next_primary, (uint64_t)64);
    }
    claim = gcq::bcast64(claim, 0);
    gcq::vm_wait();
    uint32_t gen = gcq::ld_rlx(&job->gen);
    if (claim != ~0ull) gen = (uint32_t)(claim >> 32);  // fields of that generation are already published
    const uint32_t npx = gcq::ld_rlx(&job->npx);
    const int32_t x0 = gcq::ld_rlx(&job->x0), y0 = gcq::ld_rlx(&job->y0), w = gcq::ld_rlx(&job->w);
    float* img = (float*)gcq::ld_rlx(&job->img);
    rtk::RayRec* recs = (rtk::RayRec*)gcq::ld_rlx(&job->recs);
    const uint64_t jid = gcq::ld_rlx(&job->id);
    bool have = false;
    uint32_t i = 0;
    rtk::Ray r;
    if (claim != ~0ull && (uint32_t)claim + lane < npx) {
      i = (uint32_t)claim + lane;
      r = rtk::primary(d.sc, d.W, d.H, x0 + (int32_t)(i % (uint32_t)w), y0 + (int32_t)(i / (uint32_t)w));
      have = true;
    } else {
      uint32_t item = 0;
      if (q.pop(rtid, &item)) {
        if ((item >> 16) == gen) {
          i = item & 0xFFFFu;
          r = rtk::load_ray(&recs[i]);
          have = true;
        } else if (!q.push(rtid, item)) {  // a newer job's ray: leave it for fresh fields
          gcq::or_rlx(&d.c->err, 2u);
        }
      }
    }
    if (__ballot(have) == 0) {
      j = (j + 1) % d.njobs;
      __builtin_amdgcn_s_sleep(2);
      continue;
    }
    uint32_t ended = 0;
    if (have) {
      if (rtk::shade(d.sc, d.max_depth, light, r)) {
        rtk::store_ray(&recs[i], r);
        gcq::vm_wait();  // record complete before its index is published
        if (!q.push(rtid, i | (gen << 16))) {
          gcq::or_rlx(&d.c->err, 2u);
          ended = 1;
        }
      } else {
        // Device-scope stores: this arena block may next hold another
        // request's state, written and read from other XCDs while this
        // kernel still runs; a plain store could linger dirty in this XCD's
        // L2 and be written back over it later.
        gcq::st_rlx(&img[3 * i + 0], r.acc.x);
        gcq::st_rlx(&img[3 * i + 1], r.acc.y);
        gcq::st_rlx(&img[3 * i + 2], r.acc.z);
        ended = 1;
      }
    }
    const uint64_t e = gcq::wave_sum((uint64_t)ended), t = gcq::wave_sum((uint64_t)(have ? 1u : 0u));
    if (lane == 0) {
      gcq::faa_rlx(&job->traced, t);
      gcq::vm_wait();  // every wave's traced count lands before its alive decrement
      if (e && gcq::faa_rlx(&job->alive, (uint64_t)0 - e) == e) {
        const uint64_t rays = gcq::ld_rlx(&job->traced);
        gcq::st_rlx(&job->active, 0u);
        gcq::faa_rlx(&d.c->active_rt, -1);
        gcq::vm_wait();  // slot free before the host can reuse it
        publish(d, jid, GCQ_REQ_RT, 0, (int32_t)j, rays);
      }
    }
  }
}

// Blocks [0, bfs_blocks) are the BFS lanes' sub-grids; the first wave after
// them is the gate; the waves after that run tile jobs (when the server has a
// scene).
template <class F>
__global__ void k_serve(Frontiers<F> fq0, Frontiers<F> fq1, Dev d) {
  __shared__ uint32_t stage[16][128];
  uint32_t* const st = stage[(threadIdx.x / 64) & 15];
  if (blockIdx.x < d.lane_blocks[0]) bfs_part(fq0, d, 0, 0u, st);
  else if (blockIdx.x < d.bfs_blocks) bfs_part(fq1, d, 1, d.lane_blocks[0], st);
  else if (blockIdx.x == d.bfs_blocks && threadIdx.x < 64) gate(d);
  else if (d.rt_workers) rt_part(d);
}

__global__ void k_init_rings(char* pool, uint64_t stride, uint32_t n, uint32_t nrings) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (uint32_t t = 0; t < nrings; ++t) gcq::lfq::init_slice((gcq::lfq::Ring*)(pool + t * stride), n, false, 0, gid, st, 0);
}

// First-fit allocator over the request-state arena, in 256-B units.
struct Arena {
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void init(uint64_t size) {
    free_.clear();
    free_[0] = size;
  }
  bool alloc(uint64_t n, uint64_t* off) {
    n = (n + 255) & ~255ull;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const uint64_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 255) & ~255ull;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Pending {
  gcq_request r;
  int32_t index;  // position in a trace, or -1
};
struct InFlight {
  gcq_request r;
  int32_t index, slot, lane;
  int32_t group, bit;  // multi-source member, or -1
  uint64_t off1, len1, off2, len2;
  uint32_t nv;
  double admit_ms;
};
struct Kept {
  int32_t type;
  uint64_t off, bytes;
};

}  // namespace sv

struct gcq_server {
  std::vector<gcq_graph*> graphs;
  gcq_rt_scene* scene = nullptr;
  int32_t W = 0, H = 0;
  gcq_server_config cfg{};
  hipStream_t stream = nullptr;
  gcqf::FrontierSet fs[sv::kLanes];  // per BFS lane
  uint64_t lane_cap[sv::kLanes] = {0, 0};  // frontier ring capacity per lane
  std::vector<int32_t> glane;        // the BFS lane of each graph
  uint32_t vbits = 1;
  // device
  bk::Slot* slots = nullptr;
  uint32_t* alive = nullptr;
  uint32_t* slot_active = nullptr;
  sv::GraphDev* dgraphs = nullptr;
  sv::RtJob* jobs = nullptr;
  char* rings = nullptr;
  uint64_t ring_stride = 0;
  sv::Ctl* ctl = nullptr;
  char* subq = nullptr;        // the submission queues (G-LFQ rings): BFS lane 0, BFS lane 1, RT
  uint64_t subq_off[3] = {0, 0, 0};
  sv::Req* desc = nullptr;     // request descriptors: bfs_slots BFS, then rt_slots RT
  // multi-source groups: per graph ceil(bfs_slots / 64) groups of 64 members
  bk::Group* groups = nullptr;
  sv::GroupCtl* gctl = nullptr;
  uint64_t* gdepth = nullptr;
  int32_t* gstart = nullptr;
  uint32_t* gslot = nullptr;
  uint64_t* gbits = nullptr;         // seen, nxt[0], nxt[1] of every group
  bk::HubPool* hubs = nullptr;       // per BFS lane, two level-parity pools of long adjacencies
  std::vector<std::vector<int32_t>> graph_groups;  // group indices of each graph
  std::vector<uint64_t> gbusy;       // members in flight, per group (host view)
  char* arena_base = nullptr;  // in-flight request state, budget_bytes
  sv::Arena arena;
  char* keep_base = nullptr;   // kept results, keep_bytes
  sv::Arena keep;
  // pinned
  sv::Req* sub = nullptr;
  sv::Comp* comp = nullptr;
  uint32_t* stop = nullptr;
  // host state
  bool running = false, stopped = false;
  uint64_t sub_tail = 0, comp_head = 0;
  std::vector<int32_t> free_bfs, free_rt;
  uint64_t frontier_used[sv::kLanes] = {0, 0};
  std::deque<sv::Pending> pending;
  std::map<uint64_t, sv::InFlight> inflight;  // by internal id (submission index)
  std::map<uint64_t, sv::Kept> kept;          // by caller id
  std::deque<gcq_completion> ready;
  std::vector<int32_t> ready_index;
  std::chrono::steady_clock::time_point t0;
  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

constexpr double kStopTimeoutMs = 30000.0;

void release_all(gcq_server* s) {
  for (void* p : {(void*)s->slots, (void*)s->alive, (void*)s->slot_active, (void*)s->dgraphs, (void*)s->jobs,
                  (void*)s->rings, (void*)s->ctl, (void*)s->subq, (void*)s->desc, (void*)s->arena_base,
                  (void*)s->keep_base, (void*)s->groups, (void*)s->gctl, (void*)s->gdepth, (void*)s->gstart,
                  (void*)s->gslot, (void*)s->gbits, (void*)s->hubs})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)s->sub, (void*)s->comp, (void*)s->stop})
    if (p) (void)hipHostFree(p);
  for (auto& f : s->fs) f.release();
}

uint64_t rt_bytes(int32_t w, int32_t h) { return (uint64_t)w * h * (12ull + sizeof(rtk::RayRec)); }
// A BFS query's visited bitmap: one bit per vertex, whole 128-B lines.
uint64_t vis_bytes(uint64_t nv) { return (nv + 1023) / 1024 * 128; }

int check_request(gcq_server* s, const gcq_request* r) {
  GCQ_REQUIRE(!r->keep || s->keep_base, "kept results need keep_bytes > 0");
  if (r->type == GCQ_REQ_BFS) {
    GCQ_REQUIRE(r->graph >= 0 && r->graph < (int32_t)s->graphs.size(), "bad graph %d", r->graph);
    const uint64_t nv = (uint64_t)s->graphs[r->graph]->host.nv;
    const uint64_t qbytes = 4 * nv + (s->cfg.multi_source ? 0 : vis_bytes(nv));
    GCQ_REQUIRE(qbytes <= s->cfg.budget_bytes, "query needs %llu bytes, above the budget", (unsigned long long)qbytes);
    GCQ_REQUIRE(nv <= s->lane_cap[s->glane[r->graph]], "graph larger than its lane's frontier ring");
  } else if (r->type == GCQ_REQ_RT) {
    GCQ_REQUIRE(s->scene, "server has no scene");
    GCQ_REQUIRE(r->w > 0 && r->h > 0 && (uint64_t)r->w * r->h <= sv::kMaxTilePixels, "tile must hold 1..65536 px");
    GCQ_REQUIRE(r->x0 >= 0 && r->y0 >= 0 && r->x0 + r->w <= s->W && r->y0 + r->h <= s->H, "tile outside the frame");
    GCQ_REQUIRE(rt_bytes(r->w, r->h) <= s->cfg.budget_bytes, "tile above the budget");
  } else {
    GCQ_REQUIRE(false, "unknown request type %d", r->type);
  }
  return GCQ_OK;
}

// Install pending requests, in arrival order, while a slot and memory allow.
int try_admit(gcq_server* s) {
  while (!s->pending.empty()) {
    sv::Pending& p = s->pending.front();
    const gcq_request& r = p.r;
    GCQ_REQUIRE(s->sub_tail < s->cfg.max_requests, "server request limit (%u) reached", s->cfg.max_requests);
    sv::InFlight f{};
    f.r = r;
    f.index = p.index;
    sv::Req e{};
    e.id = s->sub_tail;
    e.type = r.type;
    if (r.type == GCQ_REQ_BFS) {
      gcq_graph* g = s->graphs[r.graph];
      const uint32_t nv = (uint32_t)g->host.nv;
      const int32_t lane = s->glane[r.graph];
      const bool ms = s->cfg.multi_source != 0;
      if (s->free_bfs.empty() || (!ms && s->frontier_used[lane] + nv > s->lane_cap[lane])) break;
      int32_t gsel = -1, bsel = -1;
      if (ms) {  // a free member bit in one of the graph's groups
        for (int32_t gi : s->graph_groups[r.graph])
          if (~s->gbusy[gi]) {
            gsel = gi;
            bsel = __builtin_ctzll(~s->gbusy[gi]);
            break;
          }
        if (gsel < 0) break;
      }
      sv::Arena& a1 = r.keep ? s->keep : s->arena;
      if (!a1.alloc(4ull * nv, &f.off1)) {
        GCQ_REQUIRE(!r.keep, "keep_bytes exhausted by kept results");
        break;
      }
      const uint64_t vb = ms ? 0 : vis_bytes(nv);  // the groups' seen words replace the bitmap
      if (vb && !s->arena.alloc(vb, &f.off2)) {
        a1.release(f.off1, 4ull * nv);
        break;
      }
      f.len1 = 4ull * nv;
      f.len2 = vb;
      f.group = gsel;
      f.bit = bsel;
      if (ms) {
        s->gbusy[gsel] |= 1ull << bsel;
        e.x0 = gsel;
        e.y0 = bsel;
      }
      f.nv = nv;
      f.slot = s->free_bfs.back();
      f.lane = lane;
      s->free_bfs.pop_back();
      if (!ms) s->frontier_used[lane] += nv;  // groups' items are bounded at start
      e.graph = r.graph;
      e.src = (r.src >= 0 && r.src < (int32_t)nv) ? r.src : 0;  // bfs_bench.cpp:555
      e.buf = (uint64_t)((r.keep ? s->keep_base : s->arena_base) + f.off1);
      e.buf2 = vb ? (uint64_t)(s->arena_base + f.off2) : 0;
    } else {
      if (s->free_rt.empty()) break;
      f.group = f.bit = -1;
      const uint64_t img = 12ull * r.w * r.h, rec = (uint64_t)sizeof(rtk::RayRec) * r.w * r.h;
      sv::Arena& a1 = r.keep ? s->keep : s->arena;
      if (!a1.alloc(img, &f.off1)) {
        GCQ_REQUIRE(!r.keep, "keep_bytes exhausted by kept results");
        break;
      }
      if (!s->arena.alloc(rec, &f.off2)) {
        a1.release(f.off1, img);
        break;
      }
      f.len1 = img;
      f.len2 = rec;
      f.slot = s->free_rt.back();
      s->free_rt.pop_back();
      e.x0 = r.x0;
      e.y0 = r.y0;
      e.w = r.w;
      e.h = r.h;
      e.buf = (uint64_t)((r.keep ? s->keep_base : s->arena_base) + f.off1);
      e.buf2 = (uint64_t)(s->arena_base + f.off2);
    }
    e.slot = f.slot;
    f.admit_ms = s->now_ms();
    sv::Req* dst = &s->sub[s->sub_tail];
    std::memcpy((char*)dst + 8, (const char*)&e + 8, sizeof e - 8);
    __atomic_store_n(&dst->seq, s->sub_tail + 1, __ATOMIC_RELEASE);
    s->inflight[s->sub_tail] = f;
    ++s->sub_tail;
    s->pending.pop_front();
  }
  // Nothing in flight will free anything: the head can never be admitted.
  GCQ_REQUIRE(s->pending.empty() || !s->inflight.empty(), "request %llu does not fit the server's slots or budget",
              (unsigned long long)s->pending.front().r.id);
  return GCQ_OK;
}

void collect(gcq_server* s) {
  for (;;) {
    sv::Comp* e = &s->comp[s->comp_head];
    if (s->comp_head >= s->cfg.max_requests || __atomic_load_n(&e->seq, __ATOMIC_ACQUIRE) != s->comp_head + 1) break;
    ++s->comp_head;
    auto it = s->inflight.find(e->id);
    if (it == s->inflight.end()) continue;
    sv::InFlight& f = it->second;
    gcq_completion c{};
    c.id = f.r.id;
    c.type = f.r.type;
    c.status = e->status;
    c.levels = e->levels;
    c.rays = e->rays;
    c.arrival_ms = f.r.arrival_ms;
    c.admit_ms = f.admit_ms;
    c.done_ms = s->now_ms();
    if (f.r.type == GCQ_REQ_BFS) {
      s->free_bfs.push_back(f.slot);
      if (f.group >= 0) s->gbusy[f.group] &= ~(1ull << f.bit);
      else s->frontier_used[f.lane] -= f.nv;
    } else {
      s->free_rt.push_back(f.slot);
    }
    if (f.len2) s->arena.release(f.off2, f.len2);  // BFS bitmap or RT ray records
    if (f.r.keep) s->kept[f.r.id] = sv::Kept{f.r.type, f.off1, f.len1};
    else s->arena.release(f.off1, f.len1);  // kept results stay in the keep arena until destroy
    s->ready.push_back(c);
    s->ready_index.push_back(f.index);
    s->inflight.erase(it);
  }
}

// While requests are in flight the persistent kernel ends only after a lane's
// barrier gave up; report that instead of waiting for completions that will
// not come.
int check_kernel_alive(gcq_server* s) {
  const hipError_t q = hipStreamQuery(s->stream);
  if (q == hipErrorNotReady) return GCQ_OK;
  if (q != hipSuccess) GCQ_HIP(q);
  s->running = false;  // nothing runs on the server's memory any more
  s->stopped = true;
  gcqh::set_error("server kernel ended with requests in flight (a BFS barrier wait exceeded its spin limit)");
  return GCQ_ETIMEOUT;
}

int submit_at(gcq_server* s, const gcq_request* r, int32_t index) {
  int rc = check_request(s, r);
  if (rc) return rc;
  s->pending.push_back(sv::Pending{*r, index});
  return try_admit(s);
}

}  // namespace

extern "C" void gcq_server_config_default(gcq_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->kind = GCQ_GLFQ;
  c->bfs_threads = 65536;
  c->rt_threads = 16384;
  c->block = 256;
  c->bfs_slots = 256;
  c->rt_slots = 16;
  c->frontier_log2 = 26;
  c->max_depth = 4;
  c->pop_width = 64;
  c->flush_at = 64;
  c->budget_bytes = 64ull << 30;
  c->max_requests = 1u << 20;
}

extern "C" int gcq_server_create(gcq_graph* const* graphs, int32_t ngraphs, gcq_rt_scene* scene, int32_t frame_w,
                                 int32_t frame_h, const gcq_server_config* cfg, gcq_server** out) {
  GCQ_REQUIRE(out && cfg && (graphs || ngraphs == 0) && ngraphs >= 0, "null argument");
  GCQ_REQUIRE(ngraphs > 0 || scene, "a server needs graphs or a scene");
  GCQ_REQUIRE(cfg->block >= 64 && cfg->block <= 1024 && cfg->block % 64 == 0, "bad block");
  GCQ_REQUIRE(cfg->bfs_threads >= cfg->block && cfg->bfs_threads % cfg->block == 0, "bad bfs_threads");
  GCQ_REQUIRE(cfg->rt_threads % cfg->block == 0, "bad rt_threads");
  GCQ_REQUIRE(!scene || cfg->rt_threads >= 128, "a scene needs rt_threads >= 128 (the gate wave takes one)");
  GCQ_REQUIRE(cfg->bfs_slots >= 1 && cfg->rt_slots >= 1 && cfg->rt_slots <= 1024, "bad slot counts");
  GCQ_REQUIRE(cfg->frontier_log2 >= 10 && cfg->frontier_log2 <= 30, "frontier_log2 must be in [10, 30]");
  GCQ_REQUIRE(cfg->pop_width >= 1 && cfg->pop_width <= 64 && cfg->flush_at >= 1 && cfg->flush_at <= 64,
              "bad pop_width / flush_at");
  GCQ_REQUIRE(cfg->max_requests >= 1 && cfg->budget_bytes >= 4096 && cfg->max_depth >= 0, "bad limits");
  GCQ_REQUIRE(cfg->small_log2 <= 30, "small_log2 must be at most 30");
  GCQ_REQUIRE(cfg->multi_source <= 1, "multi_source must be 0 or 1");
  GCQ_REQUIRE(!cfg->small_log2 || (cfg->small_threads >= cfg->block && cfg->small_threads % cfg->block == 0 &&
                                   cfg->small_threads < cfg->bfs_threads),
              "the small-graph lane needs small_threads: a multiple of block below bfs_threads");
  GCQ_REQUIRE(!scene || (frame_w > 0 && frame_h > 0), "bad frame");
  gcq_server* s = new gcq_server();
  s->graphs.assign(graphs, graphs + ngraphs);
  s->scene = scene;
  s->W = frame_w;
  s->H = frame_h;
  s->cfg = *cfg;
  int32_t max_nv = 1;
  for (gcq_graph* g : s->graphs) {
    if (!g) {
      delete s;
      GCQ_REQUIRE(false, "null graph");
    }
    max_nv = std::max(max_nv, g->host.nv);
  }
  while ((1ll << s->vbits) < (int64_t)max_nv) ++s->vbits;
  // Lanes: graphs of at most 2^small_log2 vertices go to lane 1.  Each lane's
  // frontier ring holds the V sum of its in-flight queries.
  const bool two = cfg->small_log2 != 0;
  for (gcq_graph* g : s->graphs)
    s->glane.push_back(two && (uint64_t)g->host.nv <= (1ull << cfg->small_log2) ? 1 : 0);
  s->lane_cap[0] = 1ull << cfg->frontier_log2;
  s->lane_cap[1] = two ? std::min<uint64_t>(s->lane_cap[0], gcqf::pow2_at_least((uint64_t)cfg->bfs_slots
                                                                                 << cfg->small_log2))
                       : 0;
  if ((uint64_t)cfg->bfs_slots > (1ull << (31 - s->vbits))) {
    delete s;
    GCQ_REQUIRE(false, "%u BFS slots do not fit beside %u-bit vertices", cfg->bfs_slots, s->vbits);
  }
  *out = s;
  return GCQ_OK;
}

extern "C" int gcq_server_start(gcq_server* s) {
  GCQ_REQUIRE(s && !s->running && !s->stopped, "server already started");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    gcqh::set_error("no HIP device visible");
    return GCQ_ENODEV;
  }
  const gcq_server_config& c = s->cfg;
  for (gcq_graph* g : s->graphs) {
    int rc = gcqg::ensure_uploaded(g);
    if (rc) return rc;
  }
  if (s->scene) {
    int rc = rtk::scene_upload(s->scene);
    if (rc) return rc;
  }
  GCQ_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  const uint32_t lane_threads[2] = {c.bfs_threads - (c.small_log2 ? c.small_threads : 0u),
                                    c.small_log2 ? c.small_threads : 0u};
  int rc = GCQ_OK;
  for (int l = 0; l < sv::kLanes && rc == GCQ_OK; ++l)
    if (lane_threads[l]) rc = s->fs[l].ensure(c.kind, s->lane_cap[l], lane_threads[l]);
  if (rc) return rc;
  const uint32_t S = c.bfs_slots, J = c.rt_slots;
  GCQ_HIP(hipMalloc(&s->slots, sizeof(bk::Slot) * S));
  GCQ_HIP(hipMalloc(&s->alive, 4ull * bk::kAliveStride * S));
  GCQ_HIP(hipMalloc(&s->slot_active, 4ull * S));
  GCQ_HIP(hipMalloc(&s->dgraphs, sizeof(sv::GraphDev) * std::max<size_t>(1, s->graphs.size())));
  GCQ_HIP(hipMalloc(&s->jobs, sizeof(sv::RtJob) * J));
  s->ring_stride = (gcq::lfq::ring_bytes(sv::kMaxTilePixels) + 255) & ~255ull;
  GCQ_HIP(hipMalloc(&s->rings, s->ring_stride * J));
  GCQ_HIP(hipMalloc(&s->ctl, sizeof(sv::Ctl)));
  GCQ_HIP(hipMalloc(&s->hubs, 2ull * sv::kLanes * sizeof(bk::HubPool)));
  // Submission queues (BFS lane 0, BFS lane 1, RT): a slot is queued at
  // most once, so capacity = slots.
  const uint32_t qn[3] = {gcqf::pow2_at_least(std::max<uint64_t>(S, 64)), gcqf::pow2_at_least(std::max<uint64_t>(S, 64)),
                          gcqf::pow2_at_least(std::max<uint64_t>(J, 64))};
  uint64_t qbytes = 0;
  for (int q = 0; q < 3; ++q) {
    s->subq_off[q] = qbytes;
    qbytes += (gcq::lfq::ring_bytes(qn[q]) + 255) & ~255ull;
  }
  GCQ_HIP(hipMalloc(&s->subq, qbytes));
  GCQ_HIP(hipMalloc(&s->desc, sizeof(sv::Req) * (S + J)));
  // Multi-source groups: ceil(S / 64) per graph, so any S queries can be in
  // flight on one graph; seen / nxt[0] / nxt[1] words per group and vertex.
  std::vector<bk::Group> hg;
  uint64_t gwords = 0;
  if (c.multi_source) {
    const int32_t per_graph = (int32_t)((S + 63) / 64);
    s->graph_groups.assign(s->graphs.size(), {});
    uint64_t lane_items[sv::kLanes] = {0, 0};
    for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
      const uint64_t nv = (uint64_t)s->graphs[gi]->host.nv;
      for (int32_t k = 0; k < per_graph; ++k) {
        s->graph_groups[gi].push_back((int32_t)hg.size());
        bk::Group G{};
        G.off = (uint64_t)s->graphs[gi]->d_off;
        G.idx = (uint64_t)s->graphs[gi]->d_idx;
        G.seen = gwords;  // word offsets until gbits exists
        G.nxt[0] = gwords + nv;
        G.nxt[1] = gwords + 2 * nv;
        gwords += 3 * nv;
        hg.push_back(G);
        lane_items[s->glane[gi]] += nv;  // one (group, vertex) item per level at most
      }
    }
    GCQ_REQUIRE(hg.size() <= (1ull << (31 - s->vbits)), "%zu groups do not fit beside %u-bit vertices", hg.size(),
                s->vbits);
    for (int l = 0; l < sv::kLanes; ++l)
      GCQ_REQUIRE(lane_items[l] <= s->lane_cap[l], "lane %d's frontier ring (%llu) is below its groups' %llu vertices",
                  l, (unsigned long long)s->lane_cap[l], (unsigned long long)lane_items[l]);
    const uint64_t ng = hg.size();
    GCQ_HIP(hipMalloc(&s->groups, sizeof(bk::Group) * ng));
    GCQ_HIP(hipMalloc(&s->gctl, sizeof(sv::GroupCtl) * ng));
    GCQ_HIP(hipMalloc(&s->gdepth, 8ull * 64 * ng));
    GCQ_HIP(hipMalloc(&s->gstart, 4ull * 64 * ng));
    GCQ_HIP(hipMalloc(&s->gslot, 4ull * 64 * ng));
    GCQ_HIP(hipMalloc(&s->gbits, 8ull * gwords));
    std::vector<sv::GroupCtl> hc(ng);
    for (uint64_t k = 0; k < ng; ++k) {
      hg[k].seen = (uint64_t)(s->gbits + hg[k].seen);
      hg[k].nxt[0] = (uint64_t)(s->gbits + hg[k].nxt[0]);
      hg[k].nxt[1] = (uint64_t)(s->gbits + hg[k].nxt[1]);
      hg[k].depth = (uint64_t)(s->gdepth + 64 * k);
      hg[k].start = (uint64_t)(s->gstart + 64 * k);
      hg[k].fresh = (uint64_t)&s->gctl[k].fresh[0];
      hc[k] = sv::GroupCtl{};
    }
    for (size_t gi = 0; gi < s->graphs.size(); ++gi)
      for (int32_t k : s->graph_groups[gi]) {
        hc[k].lane = (uint32_t)s->glane[gi];
        hc[k].nv = (uint32_t)s->graphs[gi]->host.nv;
      }
    GCQ_HIP(hipMemcpy(s->groups, hg.data(), sizeof(bk::Group) * ng, hipMemcpyHostToDevice));
    GCQ_HIP(hipMemcpy(s->gctl, hc.data(), sizeof(sv::GroupCtl) * ng, hipMemcpyHostToDevice));
    s->gbusy.assign(ng, 0ull);
  }
  GCQ_HIP(hipMalloc(&s->arena_base, c.budget_bytes));
  s->arena.init(c.budget_bytes);
  if (c.keep_bytes) {
    GCQ_HIP(hipMalloc(&s->keep_base, c.keep_bytes));
    s->keep.init(c.keep_bytes);
  }
  GCQ_HIP(hipHostMalloc((void**)&s->sub, sizeof(sv::Req) * (size_t)c.max_requests, hipHostMallocCoherent));
  GCQ_HIP(hipHostMalloc((void**)&s->comp, sizeof(sv::Comp) * (size_t)c.max_requests, hipHostMallocCoherent));
  GCQ_HIP(hipHostMalloc((void**)&s->stop, 64, hipHostMallocCoherent));
  std::memset(s->sub, 0, sizeof(sv::Req) * (size_t)c.max_requests);
  std::memset(s->comp, 0, sizeof(sv::Comp) * (size_t)c.max_requests);
  __atomic_store_n(s->stop, 0u, __ATOMIC_RELEASE);
  std::vector<sv::GraphDev> gd(std::max<size_t>(1, s->graphs.size()));
  for (size_t i = 0; i < s->graphs.size(); ++i)
    gd[i] = sv::GraphDev{(uint64_t)s->graphs[i]->d_off, (uint64_t)s->graphs[i]->d_idx,
                         (uint32_t)s->graphs[i]->host.nv, (uint32_t)s->glane[i], 0};
  GCQ_HIP(hipMemcpy(s->dgraphs, gd.data(), sizeof(sv::GraphDev) * gd.size(), hipMemcpyHostToDevice));
  auto zero = [&](void* p, uint64_t bytes) {
    hipLaunchKernelGGL(gcqf::kf_zero_u64, dim3(gcqf::grid_of(bytes / 8)), dim3(256), 0, s->stream, (uint64_t*)p,
                       bytes / 8);
  };
  zero(s->slots, sizeof(bk::Slot) * S);
  zero(s->alive, ((4ull * bk::kAliveStride * S) + 7) & ~7ull);
  zero(s->slot_active, (4ull * S + 7) & ~7ull);
  zero(s->jobs, sizeof(sv::RtJob) * J);
  zero(s->ctl, sizeof(sv::Ctl));
  for (auto& f : s->fs)
    if (f.kind >= 0) f.reset(s->stream);
  hipLaunchKernelGGL(sv::k_init_rings, dim3(gcqf::grid_of(2ull * sv::kMaxTilePixels)), dim3(256), 0, s->stream,
                     s->rings, s->ring_stride, sv::kMaxTilePixels, J);
  for (int q = 0; q < 3; ++q)
    hipLaunchKernelGGL(sv::k_init_rings, dim3(gcqf::grid_of(2ull * qn[q])), dim3(256), 0, s->stream,
                       s->subq + s->subq_off[q], (uint64_t)0, qn[q], 1u);
  zero(s->desc, sizeof(sv::Req) * (S + J));
  if (c.multi_source) {
    zero(s->gbits, 8ull * gwords);
    zero(s->gdepth, 8ull * 64 * hg.size());
    zero(s->gstart, 4ull * 64 * hg.size());
    zero(s->gslot, 4ull * 64 * hg.size());
  }
  zero(s->hubs, 2ull * sv::kLanes * sizeof(bk::HubPool));
  GCQ_HIP(hipGetLastError());
  GCQ_HIP(hipStreamSynchronize(s->stream));

  sv::Dev d{};
  d.bd = bk::Dev{s->slots, s->alive, nullptr, &s->ctl->err, S, s->vbits, c.pop_width, c.flush_at};
  d.slots = s->slots;
  d.slot_active = s->slot_active;
  d.graphs = s->dgraphs;
  d.jobs = s->jobs;
  d.rings = s->rings;
  d.ring_stride = s->ring_stride;
  d.sub = s->sub;
  for (int q = 0; q < 3; ++q) d.subq[q] = (gcq::lfq::Ring*)(s->subq + s->subq_off[q]);
  d.desc[0] = s->desc;
  d.desc[1] = s->desc + S;
  d.comp = s->comp;
  d.stop = s->stop;
  d.c = s->ctl;
  if (s->scene) d.sc = rtk::scene_view(s->scene);
  d.W = s->W;
  d.H = s->H;
  d.max_depth = c.max_depth;
  d.njobs = J;
  d.bfs_blocks = c.bfs_threads / c.block;
  d.lane_blocks[0] = lane_threads[0] / c.block;
  d.lane_blocks[1] = lane_threads[1] / c.block;
  d.ms = c.multi_source ? 1u : 0u;
  d.ngroups = (uint32_t)hg.size();
  d.md = bk::MsDev{s->groups, &s->ctl->err, (uint32_t)hg.size(), s->vbits, c.pop_width, c.flush_at};
  d.groups = s->groups;
  d.gctl = s->gctl;
  d.gdepth = s->gdepth;
  d.gstart = s->gstart;
  d.gslot = s->gslot;
  d.hub = s->hubs;
  d.max_requests = c.max_requests;
  d.rt_workers = s->scene ? 1u : 0u;
  // The gate wave leads the first block after the BFS sub-grid; a server
  // without RT threads gets one block for it.
  const uint32_t grid = (c.bfs_threads + c.rt_threads) / c.block + (c.rt_threads == 0 ? 1u : 0u);
  int dev = 0, cus = 0, per_cu = 0;
  GCQ_HIP(hipGetDevice(&dev));
  GCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  rc = s->fs[0].dispatch([&](const auto& fq) {
    using F = std::decay_t<decltype(fq.q[0])>;
    // lane 1 uses the same queue kind; without it, its (unused) view is lane 0's
    const gcqf::Frontiers<F> fq1 = lane_threads[1] ? s->fs[1].template view<F>() : fq;
    GCQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sv::k_serve<F>, c.block, 0));
    if ((int64_t)grid > (int64_t)per_cu * cus) {
      gcqh::set_error("server grid of %u blocks does not fit resident (%d)", grid, per_cu * cus);
      return GCQ_EINVAL;
    }
    hipLaunchKernelGGL(sv::k_serve<F>, dim3(grid), dim3(c.block), 0, s->stream, fq, fq1, d);
    GCQ_HIP(hipGetLastError());
    return GCQ_OK;
  });
  if (rc) return rc;
  s->free_bfs.clear();
  for (int32_t i = (int32_t)S - 1; i >= 0; --i) s->free_bfs.push_back(i);
  s->free_rt.clear();
  for (int32_t i = (int32_t)J - 1; i >= 0; --i) s->free_rt.push_back(i);
  s->t0 = std::chrono::steady_clock::now();
  s->running = true;
  return GCQ_OK;
}

extern "C" int gcq_server_submit(gcq_server* s, const gcq_request* r) {
  GCQ_REQUIRE(s && r && s->running, "server not running");
  return submit_at(s, r, -1);
}

extern "C" int gcq_server_poll(gcq_server* s, gcq_completion* out, int32_t max, int32_t* n) {
  GCQ_REQUIRE(s && s->running, "server not running");
  collect(s);
  int rc = try_admit(s);
  int32_t k = 0;
  while (k < max && !s->ready.empty()) {
    if (out) out[k] = s->ready.front();
    s->ready.pop_front();
    s->ready_index.erase(s->ready_index.begin());
    ++k;
  }
  if (n) *n = k;
  return rc;
}

extern "C" int gcq_server_stop(gcq_server* s) {
  GCQ_REQUIRE(s, "null server");
  if (!s->running) return GCQ_OK;
  const double deadline = s->now_ms() + kStopTimeoutMs;
  while (!s->pending.empty() || !s->inflight.empty()) {
    collect(s);
    int rc = try_admit(s);
    if (rc) return rc;
    if (!s->inflight.empty() && (rc = check_kernel_alive(s)) != GCQ_OK) return rc;
    if (s->now_ms() > deadline) {
      gcqh::set_error("server: requests still in flight after %.0f ms", kStopTimeoutMs);
      return GCQ_ETIMEOUT;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  __atomic_store_n(s->stop, 1u, __ATOMIC_RELEASE);
  for (;;) {
    const hipError_t q = hipStreamQuery(s->stream);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) GCQ_HIP(q);
    if (s->now_ms() > deadline) {
      gcqh::set_error("server: persistent kernel did not end");
      return GCQ_ETIMEOUT;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  s->running = false;
  s->stopped = true;
  sv::Ctl h;
  GCQ_HIP(hipMemcpy(&h, s->ctl, sizeof h, hipMemcpyDeviceToHost));
  if (h.lane[0].bar.abort || h.lane[1].bar.abort) {
    gcqh::set_error("server aborted: a BFS barrier wait exceeded its spin limit");
    return GCQ_ETIMEOUT;
  }
  if (h.err) {
    gcqh::set_error("server kernel reported error flags 0x%x", h.err);
    return (h.err & 8u) ? GCQ_ETIMEOUT : GCQ_EFULL;
  }
  return GCQ_OK;
}

extern "C" int gcq_server_fetch(gcq_server* s, uint64_t id, void* out) {
  GCQ_REQUIRE(s && out, "null argument");
  GCQ_REQUIRE(s->stopped, "results can be fetched once the server is stopped");
  auto it = s->kept.find(id);
  GCQ_REQUIRE(it != s->kept.end(), "no kept result for request %llu", (unsigned long long)id);
  GCQ_HIP(hipMemcpy(out, s->keep_base + it->second.off, it->second.bytes, hipMemcpyDeviceToHost));
  return GCQ_OK;
}

extern "C" int gcq_server_run_trace(gcq_server* s, const gcq_request* reqs, int32_t n, gcq_completion* out) {
  GCQ_REQUIRE(s && reqs && out && n > 0, "null argument");
  for (int32_t i = 1; i < n; ++i)
    GCQ_REQUIRE(reqs[i].arrival_ms >= reqs[i - 1].arrival_ms, "trace must be sorted by arrival");
  int rc = gcq_server_start(s);
  if (rc) return rc;
  int32_t next = 0, done = 0;
  uint32_t spins = 0;
  const double deadline = reqs[n - 1].arrival_ms + kStopTimeoutMs;
  while (done < n) {
    if ((++spins & 255u) == 0 && !s->inflight.empty() && (rc = check_kernel_alive(s)) != GCQ_OK) return rc;
    const double now = s->now_ms();
    while (next < n && reqs[next].arrival_ms <= now) {
      rc = submit_at(s, &reqs[next], next);
      if (rc) return rc;
      ++next;
    }
    collect(s);
    rc = try_admit(s);
    if (rc) return rc;
    while (!s->ready.empty()) {
      out[s->ready_index.front()] = s->ready.front();
      s->ready.pop_front();
      s->ready_index.erase(s->ready_index.begin());
      ++done;
    }
    if (s->now_ms() > deadline) {
      gcqh::set_error("trace did not finish");
      return GCQ_ETIMEOUT;
    }
    if (next < n && reqs[next].arrival_ms > s->now_ms() + 0.05) std::this_thread::yield();
  }
  return gcq_server_stop(s);
}

extern "C" int gcq_server_destroy(gcq_server* s) {
  if (!s) return GCQ_OK;
  if (s->running) {
    int rc = gcq_server_stop(s);
    if (rc == GCQ_ETIMEOUT) return rc;  // the kernel still runs: do not free what it uses
  }
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  release_all(s);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GCQ_OK;
}

extern "C" int gcq_plain_run_trace(gcq_graph* const* graphs, int32_t ngraphs, gcq_rt_scene* scene, int32_t frame_w,
                                   int32_t frame_h, const gcq_bfs_config* bcfg, const gcq_rt_config* rcfg,
                                   const gcq_request* reqs, int32_t n, gcq_completion* out) {
  GCQ_REQUIRE(reqs && out && n > 0 && bcfg && rcfg, "null argument");
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
  for (int32_t i = 0; i < n; ++i) {
    const gcq_request& r = reqs[i];
    while (now() < r.arrival_ms) std::this_thread::yield();
    gcq_completion& c = out[i];
    std::memset(&c, 0, sizeof c);
    c.id = r.id;
    c.type = r.type;
    c.arrival_ms = r.arrival_ms;
    c.admit_ms = now();
    if (r.type == GCQ_REQ_BFS) {
      GCQ_REQUIRE(r.graph >= 0 && r.graph < ngraphs, "bad graph %d", r.graph);
      gcq_bfs_result res;
      int rc = gcq_bfs_run(graphs[r.graph], r.src, bcfg, nullptr, &res);
      if (rc) return rc;
      c.levels = res.levels;
    } else {
      GCQ_REQUIRE(scene, "no scene");
      gcq_rt_config rc2 = *rcfg;
      rc2.width = frame_w;
      rc2.height = frame_h;
      rc2.x0 = r.x0;
      rc2.y0 = r.y0;
      rc2.w = r.w;
      rc2.h = r.h;
      gcq_rt_result res;
      int rc = gcq_rt_render(scene, &rc2, nullptr, &res);
      if (rc) return rc;
      c.rays = res.total_rays;
    }
    c.done_ms = now();
  }
  return GCQ_OK;
}
