// hostq.cpp — ORACLE (test infrastructure only).  Host-thread std::atomic
// restatements of the reference's four queues, plus the logged-history and
// timed-throughput drivers that use them.  They serve as (a) an independent
// implementation whose histories the checker must accept and (b) the CPU
// baseline reported beside the GPU numbers.  Every RMW is seq_cst, the
// memory model the reference assumes (gwf_ring.hpp:15-17).  Each host thread
// takes its own ticket (no wave to aggregate over).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

using u64 = uint64_t;
using u32 = uint32_t;
using Clock = std::chrono::steady_clock;

struct Probe {
  u32 steps = 0;
  u32 delay_us = 0;
  void step() { ++steps; }
  void after_ticket() {
    if (delay_us) std::this_thread::sleep_for(std::chrono::microseconds(delay_us));
  }
};

inline bool cyc_behind(u32 a, u32 b) { return ((a - b) & 0xFFu) >= 0x80u; }

// ====================================================================== G-LFQ
// glf_queue.hpp:44-582 (SCQ ring of indices + fq/aq indirection).
struct LfqRing {
  static constexpr u32 BOT = 0xFFFFFFFEu, BOTC = 0xFFFFFFFFu;
  u32 n, slots;
  std::atomic<u64> tail, head;
  std::atomic<int32_t> thr;
  std::unique_ptr<std::atomic<u64>[]> e;
  static u64 pack(u32 c, bool s, u32 i) { return ((u64)(c & 0xFF) << 48) | ((u64)s << 47) | i; }
  static u32 cyc(u64 x) { return (u32)(x >> 48) & 0xFF; }
  static bool safe(u64 x) { return (x >> 47) & 1; }
  static u32 idx(u64 x) { return (u32)x; }
  u32 cyc_of(u64 t) const { return (u32)(t / slots) & 0xFF; }
  u32 slot(u64 t) const { return (u32)(t & (slots - 1)); }
  LfqRing(u32 n_, bool full) : n(n_), slots(2 * n_), e(new std::atomic<u64>[2 * n_]) {
    for (u32 i = 0; i < slots; ++i) e[i] = pack(0, true, BOT);
    head = slots;
    tail = slots;
    thr = -1;
    if (full) {
      for (u32 i = 0; i < n; ++i) e[slot(slots + i)] = pack(cyc_of(slots + i), true, i);
      tail = slots + n;
      thr = 3 * (int32_t)n - 1;
    }
  }
  void catchup(u64 t, u64 h) {
    for (int i = 0; i < 32; ++i) {
      if (tail.compare_exchange_strong(t, h)) break;
      h = head.load();
      t = tail.load();
      if (t >= h) break;
    }
  }
  void enqueue(u32 index, Probe& p) {
    for (;;) {
      const u64 T = tail.fetch_add(1);
      p.step();
      p.after_ticket();
      const u32 tc = cyc_of(T), j = slot(T);
      u64 E = e[j].load();
      for (;;) {
        if (cyc_behind(cyc(E), tc) && (safe(E) || head.load() <= T) && idx(E) >= BOT) {
          if (!e[j].compare_exchange_strong(E, pack(tc, true, index))) {
            p.step();
            continue;
          }
          if (thr.load() != 3 * (int32_t)n - 1) thr.exchange(3 * (int32_t)n - 1);
          return;
        }
        break;
      }
    }
  }
  u32 dequeue(Probe& p) {
    if (thr.load() < 0) return BOT;
    for (;;) {
      const u64 H = head.fetch_add(1);
      p.step();
      p.after_ticket();
      const u32 hc = cyc_of(H), j = slot(H);
      u64 E = e[j].load();
      for (;;) {
        const u32 ec = cyc(E), ei = idx(E);
        if (ec == hc && ei < BOT) {
          e[j].fetch_or((u64)BOTC);
          return ei;
        }
        const u64 nv = ei >= BOT ? pack(hc, safe(E), BOT) : pack(ec, false, ei);
        if (cyc_behind(ec, hc) && !e[j].compare_exchange_strong(E, nv)) {
          p.step();
          continue;
        }
        break;
      }
      const u64 T = tail.load();
      if (T <= H + 1) {
        catchup(T, H + 1);
        thr.fetch_sub(1);
        return BOT;
      }
      if (thr.fetch_sub(1) <= 0) return BOT;
    }
  }
};

struct Lfq {
  LfqRing aq, fq;
  std::unique_ptr<std::atomic<u64>[]> data;
  explicit Lfq(u32 n) : aq(n, false), fq(n, true), data(new std::atomic<u64>[n]) {}
  bool enq(u32, u64 v, Probe& p) {
    const u32 i = fq.dequeue(p);
    if (i == LfqRing::BOT) return false;
    data[i].store(v);
    aq.enqueue(i, p);
    return true;
  }
  bool deq(u32, u64* o, Probe& p) {
    const u32 i = aq.dequeue(p);
    if (i == LfqRing::BOT) return false;
    *o = data[i].load();
    fq.enqueue(i, p);
    return true;
  }
};

// ====================================================================== G-WFQ
// gwf_queue.hpp + gwf_ring.hpp: SCQ fast path, cooperative slow path,
// per-thread help countdown and the record scan of finalize_request.
struct GwfRing {
  static constexpr u32 BOT = 0xFFFFFFFEu, BOTC = 0xFFFFFFFFu;
  static constexpr u64 ENQ = 1ull << 46, SAFE = 1ull << 47;
  static constexpr u64 FIN = 1, INC = 2;
  struct Rec {
    int32_t next_check;
    u32 next_tid;
    std::atomic<u32> p2_seq1{1}, p2_seq2{0};
    std::atomic<u64> p2_local{0}, p2_cnt{0};
    std::atomic<u32> seq1{1}, seq2{0}, enqueue{0}, pending{0}, index{0};
    std::atomic<u64> localTail{0}, initTail{0}, localHead{0}, initHead{0};
  };
  u32 n, slots, B, maxthr, help_delay, enq_pat, deq_pat;
  bool ref_rule;
  u64 nulltid;
  std::atomic<u64> tail, head;
  std::atomic<int32_t> thr;
  std::unique_ptr<std::atomic<u64>[]> e;
  std::unique_ptr<Rec[]> rec;

  static u64 pack(u32 note, u32 c, bool s, bool enq, u32 i) {
    return ((u64)(note & 0xFF) << 56) | ((u64)(c & 0xFF) << 48) | ((u64)s << 47) | ((u64)enq << 46) | i;
  }
  static u32 note(u64 x) { return (u32)(x >> 56) & 0xFF; }
  static u32 cyc(u64 x) { return (u32)(x >> 48) & 0xFF; }
  static bool safe(u64 x) { return (x >> 47) & 1; }
  static bool enqb(u64 x) { return (x >> 46) & 1; }
  static u32 idx(u64 x) { return (u32)x; }
  u32 cyc_of(u64 t) const { return (u32)(t / slots) & 0xFF; }
  // Note of an entry written at cycle c outside the slow path's note advance:
  // kept as is under the reference rule (gwf_ring.hpp:404-405, 446-447), else
  // a note behind c becomes c - 1, so it never falls 128 laps behind and
  // reads as ahead (the library's rule, gwfq.hpp fresh_note).
  u32 wnote(u32 n, u32 c) const { return ref_rule || !cyc_behind(n, c) ? n : ((c - 1) & 0xFF); }
  u32 slot(u64 t) const { return (u32)(t & (slots - 1)); }
  u64 cnt(u64 g) const { return g >> B; }
  u32 gtid(u64 g) const { return (u32)(g & nulltid); }
  u64 pg(u64 c, u64 t) const { return (c << B) | t; }
  int32_t tmax() const { return 3 * (int32_t)n - 1; }

  GwfRing(u32 n_, u32 maxthr_, bool full, u32 ep, u32 dp, bool rr)
      : n(n_), slots(2 * n_), maxthr(maxthr_), help_delay(64), enq_pat(ep), deq_pat(dp), ref_rule(rr),
        e(new std::atomic<u64>[2 * n_]), rec(new Rec[maxthr_]) {
    B = 5;
    while (((1u << B) - 1u) < maxthr) ++B;
    nulltid = (1ull << B) - 1;
    for (u32 i = 0; i < slots; ++i) e[i] = pack(0xFF, 0, true, true, BOT);
    for (u32 i = 0; i < maxthr; ++i) {
      rec[i].next_check = (int32_t)help_delay;
      rec[i].next_tid = i;
    }
    tail = pg(slots, nulltid);
    head = pg(slots, nulltid);
    thr = -1;
    if (full) {
      for (u32 i = 0; i < n; ++i) e[slot(slots + i)] = pack(wnote(0xFF, cyc_of(slots + i)), cyc_of(slots + i), true, true, i);
      tail = pg(slots + n, nulltid);
      thr = tmax();
    }
  }

  void finalize(u32 me, u64 h) {  // gwf_ring.hpp:310-328
    const u64 hl = h << 2;
    u32 i = (me + 1) % maxthr;
    while (i != me) {
      const u64 lt = rec[i].localTail.load();
      if ((lt & ~3ull) == hl) {
        u64 x = hl;
        rec[i].localTail.compare_exchange_strong(x, hl | FIN);
        return;
      }
      i = (i + 1) % maxthr;
    }
  }
  void consume(u32 me, u64 h, u32 j, u64 E) {
    if (!enqb(E)) finalize(me, h);
    e[j].fetch_or(ENQ | (u64)BOTC);
  }
  void catchup(u64 t, u64 h) {
    for (int i = 0; i < 32; ++i) {
      u64 cur = tail.load();
      if (cnt(cur) >= h) break;
      if (cnt(cur) != t) {
        t = cnt(cur);
        h = cnt(head.load());
        if (t >= h) break;
        continue;
      }
      if (tail.compare_exchange_strong(cur, pg(h, gtid(cur)))) break;
      h = cnt(head.load());
      t = cnt(tail.load());
      if (t >= h) break;
    }
  }
  u64 try_enq(u32 index, Probe& p) {
    const u64 T = cnt(tail.fetch_add(1ull << B));
    p.step();
    p.after_ticket();
    const u32 tc = cyc_of(T), j = slot(T);
    u64 E = e[j].load();
    for (;;) {
      if (cyc_behind(cyc(E), tc) && (safe(E) || cnt(head.load()) <= T) && idx(E) >= BOT) {
        if (!e[j].compare_exchange_strong(E, pack(wnote(note(E), tc), tc, true, true, index))) {
          p.step();
          continue;
        }
        if (thr.load() != tmax()) thr.exchange(tmax());
        return 0;
      }
      return T;
    }
  }
  u64 try_deq(u32 me, u32* out, Probe& p) {
    const u64 H = cnt(head.fetch_add(1ull << B));
    p.step();
    p.after_ticket();
    const u32 hc = cyc_of(H), j = slot(H);
    u64 E = e[j].load();
    for (;;) {
      const u32 ec = cyc(E), ei = idx(E);
      if (ec == hc && ei < BOT) {
        consume(me, H, j, E);
        *out = ei;
        return 0;
      }
      const u64 nv = ei >= BOT ? pack(wnote(note(E), hc), hc, safe(E), true, BOT) : pack(note(E), ec, false, enqb(E), ei);
      if (cyc_behind(ec, hc) && !e[j].compare_exchange_strong(E, nv)) {
        p.step();
        continue;
      }
      break;
    }
    const u64 T = cnt(tail.load());
    if (T <= H + 1) {
      catchup(T, H + 1);
      thr.fetch_sub(1);
      *out = BOT;
      return 0;
    }
    if (thr.fetch_sub(1) <= 0) {
      *out = BOT;
      return 0;
    }
    return H;
  }
  void prepare_phase2(u32 me, std::atomic<u64>* local, u64 cl) {
    Rec& r = rec[me];
    const u32 s = r.p2_seq1.load() + 1;
    r.p2_seq1.store(s);
    r.p2_local.store((u64)(uintptr_t)local);
    r.p2_cnt.store(cl);
    r.p2_seq2.store(s);
  }
  bool slow_faa(u32 me, std::atomic<u64>& g, std::atomic<u64>* local, u64* v, bool deq, Probe& p) {
    u64 cl = 0, cr = 0;
    bool ok = false;
    do {
      p.step();
      const u64 lc = local->load();
      if (lc & FIN) {
        *v = lc;
        return false;
      }
      u64 gp = g.load();
      const u32 gt = gtid(gp);
      if (gt != (u32)nulltid) {
        Rec& h = rec[gt];
        const u32 s2 = h.p2_seq2.load();
        auto* pl = (std::atomic<u64>*)(uintptr_t)h.p2_local.load();
        const u64 pc = h.p2_cnt.load();
        const u32 s1 = h.p2_seq1.load();
        if (s1 == s2 && pl) {
          u64 x = pc | INC;
          pl->compare_exchange_strong(x, pc);
        }
        const u64 clean = pg(cnt(gp), nulltid);
        if (!g.compare_exchange_strong(gp, clean)) continue;
        gp = clean;
      }
      cr = cnt(gp);
      cl = cr << 2;
      u64 vc = *v;
      if (!local->compare_exchange_strong(vc, cl | INC)) {
        *v = vc;
        if (vc & FIN) return false;
        if (!(vc & INC)) return true;
        cl = vc & ~3ull;
        cr = cl >> 2;
      } else {
        *v = cl | INC;
      }
      prepare_phase2(me, local, cl);
      u64 ex = pg(cr, nulltid);
      ok = g.compare_exchange_strong(ex, pg(cr + 1, me));
    } while (!ok);
    if (deq) thr.fetch_sub(1);
    u64 x = cl | INC;
    local->compare_exchange_strong(x, cl);
    x = pg(cr + 1, me);
    g.compare_exchange_strong(x, pg(cr + 1, nulltid));
    *v = cl;
    return true;
  }
  bool try_enq_slow(u64 T, u32 index, Rec& r, Probe& p) {
    const u32 tc = cyc_of(T), j = slot(T);
    u64 P = e[j].load();
    for (;;) {
      p.step();
      const u32 ec = cyc(P), en = note(P), ei = idx(P);
      if (cyc_behind(ec, tc) && cyc_behind(en, tc)) {
        if ((!safe(P) && cnt(head.load()) > T) || ei < BOT) {
          if (!e[j].compare_exchange_strong(P, pack(tc, ec, safe(P), enqb(P), ei))) continue;
          return false;
        }
        const u64 prod = pack(wnote(en, tc), tc, true, false, index);
        if (!e[j].compare_exchange_strong(P, prod)) continue;
        u64 tl = T << 2;
        if (r.localTail.compare_exchange_strong(tl, (T << 2) | FIN)) {
          u64 x = prod;
          e[j].compare_exchange_strong(x, pack(wnote(en, tc), tc, true, true, index));
        }
        if (thr.load() != tmax()) thr.exchange(tmax());
        return true;
      }
      if (ec != tc) return false;
      return ref_rule ? true : ei != BOT;  // reference: gwf_ring.hpp:703-704
    }
  }
  bool try_deq_slow(u64 H, Rec& r, Probe& p) {
    const u32 hc = cyc_of(H), j = slot(H);
    u64 P = e[j].load();
    for (;;) {
      p.step();
      const u32 ec = cyc(P), en = note(P), ei = idx(P);
      if (ec == hc && ei < BOT) {
        u64 hl = H << 2;
        r.localHead.compare_exchange_strong(hl, (H << 2) | FIN);
        return true;
      }
      u32 nn = en;
      u64 val = pack(wnote(nn, hc), hc, safe(P), true, BOT);
      if (ei < BOT) {
        if (cyc_behind(ec, hc) && cyc_behind(en, hc)) {
          if (!e[j].compare_exchange_strong(P, pack(hc, ec, safe(P), enqb(P), ei))) continue;
          continue;  // re-read with the advanced note (reference takes the same second round)
        }
        val = pack(nn, ec, false, enqb(P), ei);
      }
      if (cyc_behind(ec, hc) && !e[j].compare_exchange_strong(P, val)) continue;
      const u64 T = cnt(tail.load());
      if (T <= H + 1) catchup(T, H + 1);
      if (thr.load() < 0) {
        u64 hl = H << 2;
        r.localHead.compare_exchange_strong(hl, (H << 2) | FIN);
        return true;
      }
      return false;
    }
  }
  void enqueue_slow(u32 me, u64 v, u32 index, Rec& r, Probe& p) {
    while (slow_faa(me, tail, &r.localTail, &v, false, p))
      if (try_enq_slow(v >> 2, index, r, p)) break;
  }
  void dequeue_slow(u32 me, u64 v, Rec& r, Probe& p) {
    while (slow_faa(me, head, &r.localHead, &v, true, p))
      if (try_deq_slow(v >> 2, r, p)) break;
  }
  void help_threads(u32 me, Probe& p) {  // gwf_ring.hpp:843-861
    Rec& my = rec[me];
    if (--my.next_check != 0) return;
    Rec& t = rec[my.next_tid];
    if (t.pending.load()) {
      const u32 s2 = t.seq2.load();
      const u32 isenq = t.enqueue.load();
      const u32 ix = t.index.load();
      const u64 init = isenq ? t.initTail.load() : t.initHead.load();
      if (t.seq1.load() == s2) {
        if (isenq) enqueue_slow(me, init, ix, t, p);
        else dequeue_slow(me, init, t, p);
      }
    }
    my.next_check = (int32_t)help_delay;
    my.next_tid = (my.next_tid + 1) % maxthr;
  }
  void enqueue(u32 me, u32 index, Probe& p) {
    help_threads(me, p);
    u64 T = 0;
    for (u32 c = enq_pat; --c != 0;) {
      T = try_enq(index, p);
      if (T == 0) return;
    }
    Rec& r = rec[me];
    const u32 seq = r.seq1.load();
    r.localTail.store(T << 2);
    r.initTail.store(T << 2);
    r.index.store(index);
    r.enqueue.store(1);
    r.seq2.store(seq);
    r.pending.store(1);
    enqueue_slow(me, T << 2, index, r, p);
    r.pending.store(0);
    r.seq1.store(seq + 1);
  }
  u32 dequeue(u32 me, Probe& p) {
    if (thr.load() < 0) return BOT;
    help_threads(me, p);
    u32 out = BOT;
    u64 H = 0;
    for (u32 c = deq_pat; --c != 0;) {
      H = try_deq(me, &out, p);
      if (H == 0) return out;
    }
    Rec& r = rec[me];
    const u32 seq = r.seq1.load();
    r.localHead.store(H << 2);
    r.initHead.store(H << 2);
    r.enqueue.store(0);
    r.seq2.store(seq);
    r.pending.store(1);
    dequeue_slow(me, H << 2, r, p);
    r.pending.store(0);
    r.seq1.store(seq + 1);
    const u64 h = r.localHead.load() >> 2;
    const u32 j = slot(h);
    const u64 E = e[j].load();
    if (cyc(E) == cyc_of(h) && idx(E) < BOT) {
      consume(me, h, j, E);
      return idx(E);
    }
    return BOT;
  }
};

struct Gwf {
  GwfRing aq, fq;
  std::unique_ptr<std::atomic<u64>[]> data;
  Gwf(u32 n, u32 maxthr, u32 ep, u32 dp, bool rr)
      : aq(n, maxthr, false, ep, dp, rr), fq(n, maxthr, true, ep, dp, rr), data(new std::atomic<u64>[n]) {}
  bool enq(u32 me, u64 v, Probe& p) {
    const u32 i = fq.dequeue(me, p);
    if (i == GwfRing::BOT) return false;
    data[i].store(v);
    aq.enqueue(me, i, p);
    return true;
  }
  bool deq(u32 me, u64* o, Probe& p) {
    const u32 i = aq.dequeue(me, p);
    if (i == GwfRing::BOT) return false;
    *o = data[i].load();
    fq.enqueue(me, i, p);
    return true;
  }
};

// ====================================================================== SFQ
// sfqueue_hip.cpp:89-128 with 64-bit tickets/ids/items.
struct Sfq {
  u64 qlen;
  std::atomic<u64> tail{0}, head{0};
  std::unique_ptr<std::atomic<u64>[]> ids, items;
  explicit Sfq(u64 q) : qlen(q), ids(new std::atomic<u64>[q]), items(new std::atomic<u64>[q]) {
    for (u64 i = 0; i < q; ++i) {
      ids[i] = 0;
      items[i] = 0;
    }
  }
  bool enq(u32, u64 v, Probe& p) {
    p.step();
    u64 t = tail.load();
    const u64 s = t % qlen, id = 2 * (t / qlen);
    if (ids[s].load() != id) return false;
    if (!tail.compare_exchange_strong(t, t + 1)) return false;
    items[s].store(v);
    ids[s].store(id + 1);
    return true;
  }
  bool deq_nb(u64* o, Probe& p) {
    p.step();
    u64 t = head.load();
    const u64 s = t % qlen, id = 2 * (t / qlen) + 1;
    if (ids[s].load() != id) return false;
    if (!head.compare_exchange_strong(t, t + 1)) return false;
    *o = items[s].load();
    ids[s].store(id + 1);
    return true;
  }
  // EMPTY only after observing head >= tail (bfs_bench.cpp:114-131 rule).
  bool deq(u32, u64* o, Probe& p) {
    for (;;) {
      if (deq_nb(o, p)) return true;
      const u64 h = head.load(), t = tail.load();
      if (h >= t) return false;
      std::this_thread::yield();
    }
  }
};

// ====================================================================== WFQ (YMC)
// wfqueue_hip_opt.hpp:42-568 over a flat cell array.
struct Ymc {
  static constexpr u64 BOTTOM = 0, TOP = ~0ull, EMPTY = 0, RETRY = ~1ull;
  static constexpr u32 RNULL = 0, RTOP = 1;
  struct Cell {
    std::atomic<u64> val{0};
    std::atomic<u32> enq{0}, deq{0};
  };
  struct H {
    std::atomic<u64> enq_val{0}, enq_state{0}, deq_id{0}, deq_state{0};
    u64 enq_id_last = 0;
    u32 enq_peer = 0, deq_peer = 0, next = 0;
  };
  u64 ncells;
  u32 patience;
  std::atomic<u64> T{0}, Hd{0};
  std::unique_ptr<Cell[]> cells;
  std::unique_ptr<H[]> hs;
  Ymc(u64 nc, u32 nthreads, u32 pat) : ncells(nc), patience(pat), cells(new Cell[nc]), hs(new H[nthreads]) {
    for (u32 i = 0; i < nthreads; ++i) {
      hs[i].next = (i + 1) % nthreads;
      hs[i].enq_peer = hs[i].next;
      hs[i].deq_peer = hs[i].next;
    }
  }
  static u64 ps(u32 pend, u64 id) { return ((u64)pend << 63) | (id & ((1ull << 63) - 1)); }
  static bool pend(u64 s) { return s >> 63; }
  static u64 sid(u64 s) { return s & ((1ull << 63) - 1); }
  Cell* cell(u64 i) { return i < ncells ? &cells[i] : nullptr; }
  void adv(std::atomic<u64>& e, u64 cid) {
    u64 c = e.load();
    while (c < cid && !e.compare_exchange_strong(c, cid)) {}
  }
  void commit(Cell* c, u64 v, u64 cid) {
    adv(T, cid + 1);
    c->val.exchange(v);
  }
  bool claim(std::atomic<u64>& st, u64 id, u64 cid) {
    u64 x = ps(1, id);
    return st.compare_exchange_strong(x, ps(0, cid));
  }
  bool enq_slow(u32 me, u64 v, u64 cell_id, Probe& p) {
    H& h = hs[me];
    h.enq_val.store(v);
    h.enq_state.store(ps(1, cell_id));
    for (;;) {
      p.step();
      const u64 i = T.fetch_add(1);
      Cell* c = cell(i);
      if (!c) break;
      u32 x = RNULL;
      if (c->enq.compare_exchange_strong(x, me + 2)) {
        const u64 cur = c->val.load();
        if (cur == BOTTOM || cur == TOP) {
          claim(h.enq_state, cell_id, i);
          break;
        }
      }
      if (!pend(h.enq_state.load())) break;
    }
    const u64 fid = sid(h.enq_state.load());
    Cell* d = cell(fid);
    if (!d) return false;
    // YMC's enq_slow commits its value whatever the claimed cell holds; the
    // reference commits only over TOP (wfqueue_hip_opt.hpp:308-314), which
    // loses the item once the handle's next request replaces the value.
    commit(d, v, fid);
    return true;
  }
  u64 help_enq(u32 me, Cell* c, u64 i) {
    u64 old = BOTTOM;
    if (!c->val.compare_exchange_strong(old, TOP) && old != TOP) return old;
    H& h = hs[me];
    u32 ce = c->enq.load();
    if (ce == RNULL) {
      u32 pp = h.enq_peer;
      u64 s = hs[pp].enq_state.load();
      if (h.enq_id_last != 0 && h.enq_id_last != sid(s)) {
        h.enq_id_last = 0;
        pp = hs[pp].next;
        h.enq_peer = pp;
        s = hs[pp].enq_state.load();
      }
      if (pend(s) && sid(s) > 0 && sid(s) <= i) {
        u32 prev = RNULL;
        c->enq.compare_exchange_strong(prev, pp + 2);
        if (prev != RNULL && prev != pp + 2) {
          ce = prev;
          h.enq_id_last = sid(s);
        } else {
          ce = pp + 2;
          h.enq_id_last = 0;
          h.enq_peer = hs[pp].next;
        }
      } else {
        h.enq_id_last = 0;
        h.enq_peer = hs[pp].next;
      }
      if (ce == RNULL) {
        u32 prev = RNULL;
        ce = c->enq.compare_exchange_strong(prev, RTOP) ? RTOP : prev;
      }
    }
    if (ce == RTOP) return T.load() <= i ? EMPTY : RETRY;
    H& r = hs[ce - 2];
    const u64 s = r.enq_state.load();
    const u64 v = r.enq_val.load();
    if (!pend(s)) {
      if (sid(s) == i) {
        if (c->val.load() == TOP) commit(c, v, i);
        return c->val.load();
      }
      return T.load() <= i ? EMPTY : RETRY;
    }
    if (sid(s) > i) return T.load() <= i ? EMPTY : RETRY;
    if (claim(r.enq_state, sid(s), i) || r.enq_state.load() == ps(0, i)) {
      if (c->val.load() == TOP) commit(c, v, i);
    }
    return c->val.load();
  }
  void help_deq(u32 me, u32 he, Probe& p) {
    H& x = hs[he];
    u64 s = x.deq_state.load();
    const u64 id = x.deq_id.load();
    if (!pend(s) || sid(s) < id) return;
    u64 prior = id, i = id, cand = 0;
    for (;;) {
      p.step();
      s = x.deq_state.load();
      if (!pend(s) || x.deq_id.load() != id) return;
      while (!cand && pend(s) && sid(s) == prior) {
        Cell* c = cell(++i);
        if (!c) break;
        adv(Hd, i + 1);
        const u64 v = help_enq(me, c, i);
        if (v == EMPTY || (v != TOP && v != RETRY && c->deq.load() == RNULL)) {
          cand = i;
          break;
        }
        s = x.deq_state.load();
        if (!pend(s) || x.deq_id.load() != id) return;
      }
      if (cand) {
        u64 e = ps(1, prior);
        x.deq_state.compare_exchange_strong(e, ps(1, cand));
        cand = 0;
      }
      s = x.deq_state.load();
      if (!pend(s) || x.deq_id.load() != id) return;
      const u64 a = sid(s);
      Cell* c = cell(a);
      if (!c) return;
      const u32 cur = c->deq.load();
      u32 z = RNULL;
      if (c->val.load() == TOP || c->deq.compare_exchange_strong(z, he + 2) || cur == he + 2) {
        x.deq_state.compare_exchange_strong(s, ps(0, a));
        return;
      }
      prior = a;
      if (a >= i) i = a;
    }
  }
  bool enq(u32 me, u64 v, Probe& p) {
    for (int k = (int)patience; k >= 0; --k) {
      const u64 i = T.fetch_add(1);
      p.step();
      p.after_ticket();
      Cell* c = cell(i);
      if (c) {
        u64 x = BOTTOM;
        if (c->val.compare_exchange_strong(x, v)) return true;
      }
      if (k == 0) return enq_slow(me, v, i, p);
    }
    return false;
  }
  bool deq(u32 me, u64* o, Probe& p) {
    u64 v = RETRY, cid = 0;
    for (int k = (int)patience; k >= 0; --k) {
      const u64 i = Hd.fetch_add(1);
      p.step();
      p.after_ticket();
      cid = i;
      Cell* c = cell(i);
      if (!c) {
        v = EMPTY;
        break;
      }
      const u64 r = help_enq(me, c, i);
      if (r == EMPTY) {
        v = EMPTY;
        break;
      }
      if (r != TOP && r != RETRY) {
        u32 z = RNULL;
        if (c->deq.compare_exchange_strong(z, RTOP)) {
          v = r;
          break;
        }
      }
    }
    if (v == RETRY) {  // deq_slow
      H& h = hs[me];
      h.deq_id.store(cid);
      h.deq_state.store(ps(1, cid));
      help_deq(me, me, p);
      const u64 ix = sid(h.deq_state.load());
      Cell* c = cell(ix);
      adv(Hd, ix + 1);
      v = c ? c->val.load() : EMPTY;
      if (v == TOP) v = EMPTY;
    }
    if (v != EMPTY) {
      const u32 pp = hs[me].deq_peer;
      help_deq(me, pp, p);
      hs[me].deq_peer = hs[pp].next;
      *o = v;
      return true;
    }
    return false;
  }
};

// ====================================================================== drivers
struct AnyQ {
  std::unique_ptr<Lfq> lfq;
  std::unique_ptr<Gwf> gwf;
  std::unique_ptr<Sfq> sfq;
  std::unique_ptr<Ymc> ymc;
  bool enq(u32 me, u64 v, Probe& p) {
    if (lfq) return lfq->enq(me, v, p);
    if (gwf) return gwf->enq(me, v, p);
    if (sfq) return sfq->enq(me, v, p);
    return ymc->enq(me, v, p);
  }
  bool deq(u32 me, u64* o, Probe& p, bool exact) {
    if (lfq) return lfq->deq(me, o, p);
    if (gwf) return gwf->deq(me, o, p);
    if (sfq) return exact ? sfq->deq(me, o, p) : sfq->deq_nb(o, p);
    return ymc->deq(me, o, p);
  }
};

bool make_queue(AnyQ& q, int kind, u32 cap, u32 threads, u32 ep, u32 dp, bool rr, u64 ymc_cells) {
  switch (kind) {
    case ORC_GLFQ: q.lfq.reset(new Lfq(cap)); return true;
    case ORC_GWFQ: q.gwf.reset(new Gwf(cap, std::max(threads, 1u), ep ? ep : 128, dp ? dp : 256, rr)); return true;
    case ORC_SFQ: q.sfq.reset(new Sfq(cap)); return true;
    case ORC_WFQ: q.ymc.reset(new Ymc(ymc_cells, threads, ep ? ep : 16)); return true;
  }
  return false;
}

}  // namespace

extern "C" long orc_hist_run(const orc_hist_cfg* c, orc_op* out, size_t cap, uint32_t* max_steps) {
  if (!c || c->threads == 0 || c->ops == 0) return -1;
  if ((c->capacity & (c->capacity - 1)) != 0 || c->capacity < 2) return -1;
  AnyQ q;
  const u64 ycells = (u64)c->threads * c->ops * 8 + 4096;
  if (!make_queue(q, c->kind, c->capacity, c->threads, c->enq_patience, c->deq_patience, c->ref_slow_rule != 0,
                  ycells))
    return -1;
  const u32 stride = (c->mode == 0 ? 2u : 1u) * c->ops;
  std::vector<orc_op> log((size_t)c->threads * stride);
  std::vector<u32> nlog(c->threads, 0), steps(c->threads, 0);
  std::atomic<u64> clock{1};
  std::atomic<u32> go{0};
  std::vector<std::thread> ts;
  for (u32 tid = 0; tid < c->threads; ++tid) {
    ts.emplace_back([&, tid]() {
      while (!go.load()) std::this_thread::yield();
      const bool prod = ((u64)tid * 100u) / c->threads < c->producer_percent;
      const bool victim = c->delay_every && tid % c->delay_every == 0;
      orc_op* my = &log[(size_t)tid * stride];
      u32 k = 0, mx = 0;
      for (u32 i = 0; i < c->ops; ++i) {
        if (c->mode == 0 || prod) {
          Probe p;
          p.delay_us = victim ? c->delay_us : 0;
          const u64 v = ((u64)tid << 32) | (i + 1u);
          orc_op e{tid, 0, v, 0, clock.fetch_add(1), 0};
          e.ret = q.enq(tid, v, p) ? 1 : 0;
          e.end = clock.fetch_add(1);
          my[k++] = e;
          mx = std::max(mx, p.steps);
        }
        if (c->mode == 0 || !prod) {
          Probe p;
          p.delay_us = victim ? c->delay_us : 0;
          orc_op e{tid, 1, 0, 0, clock.fetch_add(1), 0};
          u64 o = 0;
          e.ret = q.deq(tid, &o, p, true) ? o : 0;
          e.end = clock.fetch_add(1);
          my[k++] = e;
          mx = std::max(mx, p.steps);
        }
      }
      nlog[tid] = k;
      steps[tid] = mx;
    });
  }
  go.store(1);
  for (auto& t : ts) t.join();
  std::vector<orc_op> all;
  for (u32 t = 0; t < c->threads; ++t)
    for (u32 k = 0; k < nlog[t]; ++k) all.push_back(log[(size_t)t * stride + k]);
  std::sort(all.begin(), all.end(), [](const orc_op& a, const orc_op& b) { return a.call < b.call; });
  const size_t n = std::min(cap, all.size());
  if (out) std::memcpy(out, all.data(), n * sizeof(orc_op));
  if (max_steps) *max_steps = *std::max_element(steps.begin(), steps.end());
  return (long)all.size();
}

extern "C" int orc_tb_run(const orc_tb_cfg* c, orc_tb_res* r) {
  if (!c || !r || c->threads == 0) return -1;
  std::memset(r, 0, sizeof *r);
  r->hw_threads = std::thread::hardware_concurrency();
  AnyQ q;
  const double secs = (c->warmup_ms + c->run_ms) * 1e-3;
  const u64 ycells = (u64)std::min(2e8, 6e7 * secs + 1e6);
  if (!make_queue(q, c->kind, c->capacity, c->threads, 0, 0, false, ycells)) return -1;
  struct alignas(64) Cnt {
    u64 v[4] = {0, 0, 0, 0};
  };
  std::vector<Cnt> cnt(c->threads);
  std::atomic<int> phase{0};  // 0 wait, 1 warmup, 2 timed, 3 stop
  std::vector<std::thread> ts;
  for (u32 tid = 0; tid < c->threads; ++tid) {
    ts.emplace_back([&, tid]() {
      while (phase.load() == 0) std::this_thread::yield();
      const bool prod = ((u64)tid * 100u) / c->threads < c->producer_percent;
      u64 seq = 1;
      Probe p;
      int ph;
      while ((ph = phase.load(std::memory_order_relaxed)) != 3) {
        u64 l[4] = {0, 0, 0, 0};
        for (u32 i = 0; i < c->chunk_ops; ++i) {
          if (c->mode == 0 || prod) {
            const u64 v = ((u64)tid << 32) | (u32)(seq++);
            if (q.enq(tid, v, p)) ++l[0]; else ++l[1];
          }
          if (c->mode == 0 || !prod) {
            u64 o;
            if (q.deq(tid, &o, p, false)) ++l[2]; else ++l[3];
          }
        }
        if (ph == 2 && phase.load(std::memory_order_relaxed) == 2)
          for (int k = 0; k < 4; ++k) cnt[tid].v[k] += l[k];
      }
    });
  }
  phase.store(1);
  std::this_thread::sleep_for(std::chrono::milliseconds(c->warmup_ms));
  const auto t0 = Clock::now();
  phase.store(2);
  std::this_thread::sleep_for(std::chrono::milliseconds(c->run_ms));
  phase.store(3);
  const auto t1 = Clock::now();
  for (auto& t : ts) t.join();
  for (auto& x : cnt) {
    r->enq_success += x.v[0];
    r->enq_fail += x.v[1];
    r->deq_success += x.v[2];
    r->deq_empty += x.v[3];
  }
  r->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  r->succ_mops = (double)(r->enq_success + r->deq_success) / (r->elapsed_ms * 1e-3) / 1e6;
  return 0;
}
