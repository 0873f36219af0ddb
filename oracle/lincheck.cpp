// lincheck.cpp — ORACLE (test infrastructure only; never part of the product
// path).  FIFO-queue linearizability checking of operation histories.
//
// Restates the model of the reference checker correctness/lincheck/main.go,
// which drives github.com/anishathalye/porcupine v1.0.3 (a Go dependency not
// vendored in the reference and not available here):
//   JSONL record {proc, op, arg, ret, call, end}              main.go:14-21
//   op 0 = Enq(arg), ok iff ret != 0; a failed Enq is a no-op  main.go:63-66, 100-103
//   op 1 = Deq, ret 0 = Empty                                  main.go:67-75
//   Deq on empty model must be Empty; else must return head   main.go:106-116
//   check with a 10 s budget, result Ok / Illegal / Unknown    main.go:138-147
//
// Two engines:
//  * wgl  — exact Wing-Gong/Lowe search with memoisation of (linearised set,
//           queue state), the algorithm Porcupine implements.
//  * pattern — polynomial check for histories whose enqueued values are
//           unique (the case for every history this build emits): a complete
//           history of a FIFO queue with unique values is linearizable iff it
//           has none of the four bad patterns of Henzinger, Sezgin, Vafeiadis,
//           "Aspect-Oriented Linearizability Proofs" (CONCUR 2013):
//             fresh   a Deq returns a value never enqueued, or before its Enq began
//             repeat  a value is dequeued twice
//             order   Enq(a) precedes Enq(b), b is dequeued, and Deq(a) does
//                     not start before Deq(b) ends (or a is never dequeued)
//             witness an Empty Deq whose whole interval is covered by the
//                     union of "a is surely in the queue" intervals
//                     (Enq(a).end, Deq(a).call or +inf)
//           O(n log n); used where WGL would exceed Porcupine's budget.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "oracle.h"

namespace {

using Clock = std::chrono::steady_clock;

// ------------------------------------------------------------------ WGL
struct Entry {
  bool is_call;
  int id;
  uint64_t time;
  Entry* match = nullptr;  // call -> its return
  Entry* prev = nullptr;
  Entry* next = nullptr;
};

struct StateKey {
  std::vector<uint64_t> lin;   // bitset of linearised ops
  std::vector<uint64_t> queue; // model state
  bool operator==(const StateKey& o) const { return lin == o.lin && queue == o.queue; }
};
struct StateHash {
  size_t operator()(const StateKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t w : k.lin) h = (h ^ w) * 1099511628211ull;
    h ^= 0x9e3779b97f4a7c15ull;
    for (uint64_t w : k.queue) h = (h ^ w) * 1099511628211ull;
    return (size_t)h;
  }
};

// Model step (main.go:93-121).  Returns false if the output is illegal.
bool step(std::deque<uint64_t>& q, const orc_op& op) {
  if (op.op == 0) {
    if (op.ret != 0) q.push_back(op.arg);
    return true;
  }
  if (q.empty()) return op.ret == 0;
  if (op.ret == 0) return false;
  if (q.front() != op.ret) return false;
  q.pop_front();
  return true;
}

void lift(Entry* c) {
  c->prev->next = c->next;
  if (c->next) c->next->prev = c->prev;
  Entry* r = c->match;
  r->prev->next = r->next;
  if (r->next) r->next->prev = r->prev;
}
void unlift(Entry* c) {
  Entry* r = c->match;
  r->prev->next = r;
  if (r->next) r->next->prev = r;
  c->prev->next = c;
  if (c->next) c->next->prev = c;
}

int check_wgl(const orc_op* ops, size_t n, double budget_s, std::string* why) {
  if (n == 0) return ORC_OK;
  std::vector<Entry> ent(2 * n);
  std::vector<Entry*> order(2 * n);
  for (size_t i = 0; i < n; ++i) {
    ent[2 * i] = Entry{true, (int)i, ops[i].call};
    ent[2 * i + 1] = Entry{false, (int)i, ops[i].end};
    ent[2 * i].match = &ent[2 * i + 1];
    order[2 * i] = &ent[2 * i];
    order[2 * i + 1] = &ent[2 * i + 1];
  }
  std::stable_sort(order.begin(), order.end(), [](const Entry* a, const Entry* b) {
    if (a->time != b->time) return a->time < b->time;
    return a->is_call && !b->is_call;  // a call at time t precedes a return at t
  });
  Entry head{false, -1, 0};
  Entry* prev = &head;
  for (Entry* e : order) {
    prev->next = e;
    e->prev = prev;
    prev = e;
  }
  prev->next = nullptr;

  const size_t words = (n + 63) / 64;
  std::vector<uint64_t> lin(words, 0);
  std::deque<uint64_t> state;
  std::unordered_set<StateKey, StateHash> cache;
  struct Frame {
    Entry* e;
    std::deque<uint64_t> state;
  };
  std::vector<Frame> stack;
  const auto t0 = Clock::now();
  size_t iter = 0;
  Entry* e = head.next;
  while (head.next != nullptr) {
    if ((++iter & 0xFFFF) == 0 &&
        std::chrono::duration<double>(Clock::now() - t0).count() > budget_s) {
      if (why) *why = "wgl: budget exhausted";
      return ORC_UNKNOWN;
    }
    if (e->is_call) {
      std::deque<uint64_t> ns = state;
      if (step(ns, ops[e->id])) {
        std::vector<uint64_t> nl = lin;
        nl[e->id / 64] |= 1ull << (e->id % 64);
        StateKey key{nl, std::vector<uint64_t>(ns.begin(), ns.end())};
        if (cache.insert(std::move(key)).second) {
          stack.push_back(Frame{e, state});
          state.swap(ns);
          lin.swap(nl);
          lift(e);
          e = head.next;
          continue;
        }
      }
      e = e->next;
    } else {
      if (stack.empty()) {
        if (why) *why = "wgl: no linearization (operation " + std::to_string(e->id) + " cannot be placed)";
        return ORC_ILLEGAL;
      }
      Frame f = std::move(stack.back());
      stack.pop_back();
      lin[f.e->id / 64] &= ~(1ull << (f.e->id % 64));
      state.swap(f.state);
      unlift(f.e);
      e = f.e->next;
    }
  }
  return ORC_OK;
}

// ------------------------------------------------------------------ patterns
int check_patterns(const orc_op* ops, size_t n, std::string* why) {
  struct Life {
    uint64_t enq_call = 0, enq_end = 0;
    bool enqd = false;
    uint64_t deq_call = UINT64_MAX, deq_end = UINT64_MAX;
    bool deqd = false;
  };
  std::unordered_map<uint64_t, Life> life;
  life.reserve(n);
  auto fail = [&](const std::string& s) {
    if (why) *why = s;
    return ORC_ILLEGAL;
  };
  for (size_t i = 0; i < n; ++i) {
    const orc_op& o = ops[i];
    if (o.op == 0 && o.ret != 0) {
      Life& l = life[o.arg];
      if (l.enqd) {
        if (why) *why = "pattern: enqueued values are not unique";
        return ORC_UNKNOWN;
      }
      l.enqd = true;
      l.enq_call = o.call;
      l.enq_end = o.end;
    }
  }
  std::vector<size_t> empties;
  for (size_t i = 0; i < n; ++i) {
    const orc_op& o = ops[i];
    if (o.op != 1) continue;
    if (o.ret == 0) {
      empties.push_back(i);
      continue;
    }
    auto it = life.find(o.ret);
    if (it == life.end() || !it->second.enqd)
      return fail("fresh: dequeue returned " + std::to_string(o.ret) + " which no enqueue inserted");
    Life& l = it->second;
    if (o.end < l.enq_call)
      return fail("fresh: value " + std::to_string(o.ret) + " dequeued before its enqueue began");
    if (l.deqd) return fail("repeat: value " + std::to_string(o.ret) + " dequeued twice");
    l.deqd = true;
    l.deq_call = o.call;
    l.deq_end = o.end;
  }
  // order: enqueues sorted by end time, prefix max of Deq(a).call (inf if none).
  std::vector<const Life*> by_end;
  by_end.reserve(life.size());
  for (auto& kv : life)
    if (kv.second.enqd) by_end.push_back(&kv.second);
  std::sort(by_end.begin(), by_end.end(), [](const Life* a, const Life* b) { return a->enq_end < b->enq_end; });
  std::vector<uint64_t> ends(by_end.size()), pmax(by_end.size());
  uint64_t m = 0;
  for (size_t i = 0; i < by_end.size(); ++i) {
    ends[i] = by_end[i]->enq_end;
    m = std::max(m, by_end[i]->deqd ? by_end[i]->deq_call : UINT64_MAX);
    pmax[i] = m;
  }
  for (auto& kv : life) {
    const Life& b = kv.second;
    if (!b.enqd || !b.deqd) continue;
    // enqueues a with Enq(a).end < Enq(b).call
    const size_t k = std::lower_bound(ends.begin(), ends.end(), b.enq_call) - ends.begin();
    if (k > 0 && pmax[k - 1] > b.deq_end)
      return fail("order: value " + std::to_string(kv.first) +
                  " dequeued while an earlier-enqueued value was still queued");
  }
  // witness: union of surely-present intervals (Enq.end, Deq.call or inf).
  std::vector<std::pair<uint64_t, uint64_t>> iv;
  for (auto& kv : life)
    if (kv.second.enqd) iv.push_back({kv.second.enq_end, kv.second.deqd ? kv.second.deq_call : UINT64_MAX});
  std::sort(iv.begin(), iv.end());
  std::vector<std::pair<uint64_t, uint64_t>> merged;
  for (auto& p : iv) {
    if (p.second <= p.first) continue;  // empty open interval
    if (!merged.empty() && p.first < merged.back().second) {
      merged.back().second = std::max(merged.back().second, p.second);
    } else {
      merged.push_back(p);
    }
  }
  for (size_t i : empties) {
    const orc_op& d = ops[i];
    // merged interval with start < d.call, the last such one
    auto it = std::lower_bound(merged.begin(), merged.end(), std::make_pair(d.call, (uint64_t)0));
    if (it == merged.begin()) continue;
    --it;
    if (it->first < d.call && it->second >= d.end)
      return fail("witness: dequeue by proc " + std::to_string(d.proc) + " returned empty at [" +
                  std::to_string(d.call) + "," + std::to_string(d.end) + "] while the queue was surely non-empty");
  }
  return ORC_OK;
}

// ------------------------------------------------------------------ JSONL
bool field(const char* line, const char* key, uint64_t* out) {
  const char* p = strstr(line, key);
  if (!p) return false;
  p += strlen(key);
  while (*p == ' ' || *p == ':' || *p == '"') ++p;
  if (*p < '0' || *p > '9') return false;
  *out = strtoull(p, nullptr, 10);
  return true;
}

}  // namespace

extern "C" int orc_lincheck(const orc_op* ops, size_t n, int engine, double budget_s, char* msg, size_t msg_len) {
  std::string why;
  int rc;
  if (engine == ORC_ENGINE_WGL) {
    rc = check_wgl(ops, n, budget_s, &why);
  } else if (engine == ORC_ENGINE_PATTERN) {
    rc = check_patterns(ops, n, &why);
  } else {  // auto: patterns when values are unique, else WGL
    rc = check_patterns(ops, n, &why);
    if (rc == ORC_UNKNOWN) rc = check_wgl(ops, n, budget_s, &why);
  }
  if (msg && msg_len) snprintf(msg, msg_len, "%s", why.c_str());
  return rc;
}

extern "C" long orc_read_jsonl(const char* path, orc_op* out, size_t cap) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  char line[1024];
  size_t n = 0;
  while (fgets(line, sizeof line, f)) {
    uint64_t proc, op, arg, ret, call, end;
    if (!field(line, "\"proc\"", &proc) || !field(line, "\"op\"", &op) || !field(line, "\"arg\"", &arg) ||
        !field(line, "\"ret\"", &ret) || !field(line, "\"call\"", &call) || !field(line, "\"end\"", &end)) {
      if (line[0] == '\n' || line[0] == 0) continue;
      fclose(f);
      return -2;
    }
    if (n < cap) out[n] = orc_op{(uint32_t)proc, (uint32_t)op, arg, ret, call, end};
    ++n;
  }
  fclose(f);
  return (long)n;
}
