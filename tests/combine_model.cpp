// CPU model of the workgroup claim-combining protocol of
// gcq_amd/csrc/gcq/prim.hpp (combine_claim), test infrastructure only.
//
// Threads stand for the waves of one workgroup; std::atomic words stand for
// the LDS words and for the queue's Head/Tail counters.  A producer thread
// advances Tail.  Checked for every trial:
//  - every ticket handed out is unique, and the tickets form exactly
//    [0, Head): no ticket is lost or handed out twice;
//  - every member of a group started its operation before the leader read
//    Tail (an event counter taken at each wave's call and just before the
//    leader's Tail read), so an EMPTY answer is a witness inside each
//    member's operation;
//  - no wave waits forever (a spin bound turns a deadlock into a failure).
// Built with -DPLANT_LATE_CLOSE the leader reads Tail before it closes its
// group, the mistake the ordering rule exists to prevent; the member check
// must then fail.
// With a third argument "faa" the leader instead adds the group's whole lane
// count, unbounded (group_faa, the producers' Tail tickets): every lane then
// gets a ticket, and the tickets must still form exactly [0, Head).
// With a third argument "wrap" a scripted run wraps the 16-bit group id: wave
// 0 joins a group that wave 1 leads, then leads 65,535 groups alone, then
// joins wave 1's next group, whose id equals the first one's modulo 2^16.
// Built with -DPLANT_STALE_GROUP a wave does not clear its result word before
// it joins; it then takes the first group's stale result and the ticket
// checks must fail.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Combine {
  std::atomic<uint64_t> join;
  std::atomic<uint32_t> done;
  std::atomic<uint32_t> res_g[16];
  uint32_t res_k[16];
  uint64_t res_base[16];
};

std::atomic<uint64_t> g_head{0}, g_tail{0}, g_event{0};
std::atomic<uint64_t> g_call_event[16];
std::atomic<uint64_t> g_failures{0};
std::atomic<uint32_t> g_hold_for{0};  // scripted runs: waves a leader waits for before its close
bool g_plain_faa = false;
constexpr uint64_t kSpinLimit = 1ull << 28;

void fail(const char* what) {
  if (g_failures.fetch_add(1) < 5) std::fprintf(stderr, "FAIL: %s\n", what);
}

// One claim by wave w for n lanes; out[r] = ticket, or ~0 for none.
// Returns true when the group answered EMPTY.
bool claim(Combine* c, uint32_t w, uint32_t n, uint64_t* out, std::mt19937& rng) {
  g_call_event[w].store(g_event.fetch_add(1));
#ifndef PLANT_STALE_GROUP
  c->res_g[w].store(~0u, std::memory_order_relaxed);
#endif
  const uint64_t old = c->join.fetch_add((1ull << (32 + w)) + n, std::memory_order_release);
  const uint32_t grp = (uint32_t)(old >> 48);
  const uint32_t off = (uint32_t)old;
  uint32_t k = 0;
  uint64_t base = 0;
  if (off == 0) {
    for (uint64_t s = 0; (c->done.load(std::memory_order_acquire) & 0xFFFFu) != ((grp - 1u) & 0xFFFFu); ++s) {
      if (s > kSpinLimit) { fail("leader waited forever"); break; }
      std::this_thread::yield();
    }
    for (uint32_t hold = g_hold_for.load(); hold && ((c->join.load() >> 32) & hold) != hold;)
      std::this_thread::yield();
    const uint64_t h = g_head.load();
#ifdef PLANT_LATE_CLOSE
    const uint64_t tread = g_event.fetch_add(1);
    const uint64_t t = g_tail.load();
    for (int i = 0; i < 8; ++i) std::this_thread::yield();
    const uint64_t fin = c->join.exchange((uint64_t)((grp + 1u) & 0xFFFFu) << 48);
#else
    const uint64_t fin = c->join.exchange((uint64_t)((grp + 1u) & 0xFFFFu) << 48);
    if (rng() % 4 == 0) std::this_thread::yield();
    const uint64_t tread = g_event.fetch_add(1);
    const uint64_t t = g_tail.load();
#endif
    const uint32_t lanes = (uint32_t)fin, members = (uint32_t)(fin >> 32) & 0xFFFFu;
    if (!(members & (1u << w))) fail("leader is not a member of its group");
    for (uint32_t b = members; b; b &= b - 1)
      if (g_call_event[__builtin_ctz(b)].load() >= tread) fail("a member started after the leader read Tail");
    const uint64_t avail = t > h ? t - h : 0;
    k = g_plain_faa ? lanes : (avail < lanes ? (uint32_t)avail : lanes);
    if (k) base = g_head.fetch_add(k);
    for (uint32_t b = members & ~(1u << w); b; b &= b - 1) {
      const uint32_t x = (uint32_t)__builtin_ctz(b);
      c->res_k[x] = k;
      c->res_base[x] = base;
      c->res_g[x].store(grp, std::memory_order_release);
    }
    c->done.store(grp, std::memory_order_release);
  } else {
    for (uint64_t s = 0; c->res_g[w].load(std::memory_order_acquire) != grp; ++s) {
      if (s > kSpinLimit) { fail("member waited forever"); break; }
      std::this_thread::yield();
    }
    k = c->res_k[w];
    base = c->res_base[w];
  }
  for (uint32_t r = 0; r < n; ++r) out[r] = off + r < k ? base + off + r : ~0ull;
  return k == 0;
}

// Scripted 16-bit group-id wrap (see the header).  Returns the failure count.
uint64_t wrap_run() {
  g_head = 0;
  g_tail = 1ull << 50;  // every claim finds items
  Combine c;
  c.join = 1ull << 48;
  c.done = 0;
  for (auto& x : c.res_g) x = ~0u;
  std::mt19937 rng(7);
  std::vector<uint64_t> got;
  uint64_t out[64];
  auto keep = [&](uint32_t n) {
    for (uint32_t j = 0; j < n; ++j)
      if (out[j] != ~0ull) got.push_back(out[j]);
  };
  auto pair = [&]() {  // wave 1 leads, wave 0 joins as a member
    g_hold_for = 1u;
    uint64_t out1[64];
    std::mt19937 r1(11);
    std::thread lead([&] { claim(&c, 1, 5, out1, r1); });
    while (!((c.join.load() >> 33) & 1u)) std::this_thread::yield();
    claim(&c, 0, 3, out, rng);
    lead.join();
    g_hold_for = 0;
    keep(3);
    for (uint32_t j = 0; j < 5; ++j)
      if (out1[j] != ~0ull) got.push_back(out1[j]);
  };
  pair();
  for (int i = 0; i < 65535; ++i) {
    claim(&c, 0, 2, out, rng);
    keep(2);
  }
  pair();
  std::vector<uint8_t> seen(g_head.load(), 0);
  for (uint64_t t : got) {
    if (t >= seen.size()) fail("ticket at or beyond Head");
    else if (seen[t]++) fail("ticket handed out twice");
  }
  if (got.size() != g_head.load()) fail("tickets below Head not all handed out");
  std::printf("wrap run: %zu tickets, Head %llu, failures %llu\n", got.size(), (unsigned long long)g_head.load(),
              (unsigned long long)g_failures.load());
  return g_failures.load();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 3 && std::string(argv[3]) == "wrap") return wrap_run() ? 1 : 0;
  const int trials = argc > 1 ? std::atoi(argv[1]) : 30;
  const int ops = argc > 2 ? std::atoi(argv[2]) : 2000;
  g_plain_faa = argc > 3 && std::string(argv[3]) == "faa";
  uint64_t empties = 0, groups_with_items = 0;
  for (int trial = 0; trial < trials; ++trial) {
    g_head = 0;
    g_tail = 0;
    Combine c;
    c.join = 1ull << 48;
    c.done = 0;
    for (auto& x : c.res_g) x = ~0u;
    const int waves = 2 + trial % 15;
    std::vector<std::vector<uint64_t>> got(waves);
    std::atomic<bool> stop{false};
    std::thread producer([&] {
      std::mt19937 r(trial);
      while (!stop) {
        g_tail.fetch_add(r() % 96);
        for (int i = 0; i < 40; ++i) std::this_thread::yield();
      }
    });
    std::vector<std::thread> th;
    std::atomic<uint64_t> trial_empties{0};
    for (int w = 0; w < waves; ++w)
      th.emplace_back([&, w] {
        std::mt19937 rng(trial * 131 + w);
        uint64_t out[64];
        for (int i = 0; i < ops; ++i) {
          const uint32_t n = 1 + rng() % 64;
          if (claim(&c, (uint32_t)w, n, out, rng)) trial_empties++;
          for (uint32_t j = 0; j < n; ++j) {
            if (out[j] != ~0ull) got[w].push_back(out[j]);
            else if (g_plain_faa) fail("a lane got no ticket from a plain group FAA");
          }
        }
      });
    for (auto& t : th) t.join();
    stop = true;
    producer.join();
    std::vector<uint8_t> seen(g_head.load(), 0);
    uint64_t total = 0;
    for (auto& v : got)
      for (uint64_t t : v) {
        ++total;
        if (t >= seen.size()) fail("ticket at or beyond Head");
        else if (seen[t]++) fail("ticket handed out twice");
      }
    if (total != g_head.load()) fail("tickets below Head not all handed out");
    empties += trial_empties;
    groups_with_items += total ? 1 : 0;
  }
  std::printf("trials %d empties %llu failures %llu\n", trials, (unsigned long long)empties,
              (unsigned long long)g_failures.load());
  return g_failures.load() ? 1 : 0;
}
