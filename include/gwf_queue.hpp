// gwf_queue.hpp — drop-in device API of the G-WFQ wait-free queue for the
// reference harnesses (throughout_bench.cpp:112-142, correctness/history.cpp:
// 48-63, concurrent_bfs/bfs_bench.cpp:34-46, concurrent_rt/rt_final.cpp:91-135),
// backed by this library's gfx950 ring (gcq_amd/csrc/gcq/gwfq.hpp).
//
// Same names, template parameters and return conventions as the reference's
// gwf_queue.hpp:31-103 / :313-345 and gwf_ring.hpp:52-63, :229-298,
// :1234-1380: wf_config<B> and the gwf_1k..gwf_64k presets, wf_t<Cfg,N> and
// wf_mpmc_t<Cfg,N> (objects the host allocates with sizeof and a
// single-thread init kernel initialises), wf_init / wf_mpmc_init,
// Enqueue_wf / Dequeue_wf (index 0..N-1, wf_entry::BOT when empty),
// Enqueue_Ptr / Dequeue_Ptr (false = full / empty), wf_tid().  The object
// layout is this library's (counters on their own 128-B lines, records
// indexed by thread id), not the reference's; harnesses never look inside.
// GWF_N is deliberately not defined here: harnesses define it themselves.
#ifndef GCQ_DROPIN_GWF_QUEUE_HPP
#define GCQ_DROPIN_GWF_QUEUE_HPP
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gcq/gwfq.hpp"

#ifndef WF_ENQ_PATIENCE
#define WF_ENQ_PATIENCE 128
#endif
#ifndef WF_DEQ_PATIENCE
#define WF_DEQ_PATIENCE 256
#endif
#ifndef WF_HELP_DELAY
#define WF_HELP_DELAY 64
#endif

// Thread-index width of the reference's global Tail/Head word; here it
// fixes MAX_THREADS (thread ids 0 .. MAX_THREADS-1 may use a queue).
template <int THRIDX_BITS = 14>
struct wf_config {
  static_assert(THRIDX_BITS >= 5 && THRIDX_BITS <= 20, "THRIDX_BITS must be in [5, 20]");
  static constexpr int COUNTER_BITS = 64 - THRIDX_BITS;
  static constexpr uint32_t MAX_THREADS = (1u << THRIDX_BITS) - 1;
  static constexpr uint64_t NULL_TID = (1ull << THRIDX_BITS) - 1;
};
using gwf_1k = wf_config<10>;
using gwf_2k = wf_config<11>;
using gwf_4k = wf_config<12>;
using gwf_8k = wf_config<13>;
using gwf_16k = wf_config<14>;
using gwf_32k = wf_config<15>;
using gwf_64k = wf_config<16>;

struct wf_entry {
  static constexpr uint32_t BOT = 0xFFFFFFFEu;  // Dequeue_wf's "empty" (gwf_queue.hpp:162)
};

template <typename Config, uint32_t N_PARAM>
struct alignas(256) wf_t {
  static constexpr uint32_t N = N_PARAM;
  static constexpr uint32_t SLOTS = 2 * N;
  static_assert((N & (N - 1)) == 0 && N >= 64, "N must be a power of two >= 64");
  static_assert(N <= (1u << 25), "ring indices are 26-bit here: N <= 2^25");
  static_assert(Config::MAX_THREADS <= N, "wf requires k <= n (gwf_queue.hpp:320)");
  unsigned char bytes[gcq::gwf::ring_bytes(N, Config::MAX_THREADS)];
};

template <typename Config, uint32_t N_PARAM>
struct wf_mpmc_t {
  using queue_type = wf_t<Config, N_PARAM>;
  queue_type aq;             // published indices
  queue_type fq;             // free indices
  uint64_t data[N_PARAM];    // payloads
};

// 1-D global thread id, blockIdx.y ignored (gwf_ring.hpp:52-63); must be
// below Config::MAX_THREADS and unique among concurrent users of a queue.
__device__ __forceinline__ uint32_t wf_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

namespace gcq_dropin {
template <typename Config, uint32_t N>
__host__ __device__ constexpr gcq::gwf::InitParams gwf_params() {
  return gcq::gwf::InitParams{N, Config::MAX_THREADS, 0u, WF_HELP_DELAY, WF_ENQ_PATIENCE, WF_DEQ_PATIENCE, 0u};
}
template <typename Config, uint32_t N>
__device__ __forceinline__ gcq::gwf::Mpmc gwf_view(wf_mpmc_t<Config, N>* q) {
  return gcq::gwf::Mpmc{(gcq::gwf::Ring*)&q->aq, (gcq::gwf::Ring*)&q->fq, q->data};
}
}  // namespace gcq_dropin

// Initialise an empty ring; called by one thread (the harness's init kernel).
template <typename Config, uint32_t N_PARAM>
__device__ void wf_init(wf_t<Config, N_PARAM>* q) {
  gcq::gwf::init_slice((gcq::gwf::Ring*)q, gcq_dropin::gwf_params<Config, N_PARAM>(), false, 0u, 1u);
}

// aq empty, fq holding every index 0..N-1 (gwf_ring.hpp:258-298).
template <typename Config, uint32_t N_PARAM>
__device__ void wf_mpmc_init(wf_mpmc_t<Config, N_PARAM>* q) {
  const gcq::gwf::InitParams p = gcq_dropin::gwf_params<Config, N_PARAM>();
  gcq::gwf::init_slice((gcq::gwf::Ring*)&q->aq, p, false, 0u, 1u);
  gcq::gwf::init_slice((gcq::gwf::Ring*)&q->fq, p, true, 0u, 1u);
}

template <typename Config, uint32_t N_PARAM>
__device__ void Enqueue_wf(wf_t<Config, N_PARAM>* q, uint32_t index) {
  gcq::gwf::NoProbe np;
  gcq::gwf::RingOps<>((gcq::gwf::Ring*)q, wf_tid(), &np).enqueue(index);
}

template <typename Config, uint32_t N_PARAM>
__device__ uint32_t Dequeue_wf(wf_t<Config, N_PARAM>* q) {
  gcq::gwf::NoProbe np;
  const uint32_t x = gcq::gwf::RingOps<>((gcq::gwf::Ring*)q, wf_tid(), &np).dequeue(true);
  if (x == gcq::gwf::kBot) return wf_entry::BOT;
  gcq::dropin_acquire();
  return x;
}

template <typename Config, uint32_t N_PARAM>
__device__ bool Enqueue_Ptr(wf_mpmc_t<Config, N_PARAM>* q, uint64_t value) {
  return gcq::gwf::enqueue_ptr(gcq_dropin::gwf_view(q), wf_tid(), value);
}

template <typename Config, uint32_t N_PARAM>
__device__ bool Dequeue_Ptr(wf_mpmc_t<Config, N_PARAM>* q, uint64_t* out) {
  // The reference fences after taking the value (gwf_ring.hpp:1377), which
  // also orders the caller's later loads after the dequeue.
  if (!gcq::gwf::dequeue_ptr(gcq_dropin::gwf_view(q), wf_tid(), out)) return false;
  gcq::dropin_acquire();
  return true;
}

#endif  // GCQ_DROPIN_GWF_QUEUE_HPP
