// queue_impl.hpp — the gcq_queue handle and the per-kind device adapters
// that every libgcq kernel (throughput, history, FIFO check, BFS, RT)
// instantiates.  An adapter is a small by-value struct holding device
// pointers; enq/deq follow the harness adapter conventions of the reference
// (throughout_bench.cpp:112-292): enqueue returns false when full, dequeue
// returns false when empty.
#pragma once
#include "common.hpp"
#include "gcq/glfq.hpp"
#include "gcq/gwfq.hpp"
#include "gcq/sfq.hpp"
#include "gcq/ymc.hpp"

struct gcq_queue {
  gcq_queue_config cfg;
  void* dev = nullptr;      // queue object in HBM
  uint64_t bytes = 0;
  hipStream_t stream = nullptr;
  void* aux = nullptr;      // WFQ: per-thread handles
  uint64_t aux_bytes = 0;
};

namespace gcqk {

// deq's Combine argument: the kernel's LDS context for combining the
// consumer-side claims of a workgroup's idle waves.  enq's: a second LDS
// context for drawing the producers' Tail tickets together (group_faa).
// G-WFQ and G-LFQ use both, WFQ only the second; SFQ ignores them.

// begin(tid) / end(tid): called by every wave of a kernel before its first
// and after its last operation on the queue (WFQ's ring mode publishes and
// withdraws a floor per wave; the other queues need nothing).
#define GCQ_NO_LIFETIME                                  \
  __device__ __forceinline__ void begin(uint32_t) const {} \
  __device__ __forceinline__ void end(uint32_t) const {}

struct GwfQ {
  gcq::gwf::Mpmc m;
  GCQ_NO_LIFETIME
  template <class P = gcq::gwf::NoProbe>
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, P* p = nullptr, gcq::Combine* ct = nullptr) const {
    return gcq::gwf::enqueue_ptr<P>(m, tid, v, p, ct);
  }
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, decltype(nullptr), gcq::Combine* ct) const {
    return enq(tid, v, (gcq::gwf::NoProbe*)nullptr, ct);
  }
  template <class P = gcq::gwf::NoProbe>
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, P* p = nullptr, gcq::Combine* cb = nullptr) const {
    return gcq::gwf::dequeue_ptr<P>(m, tid, o, p, cb);
  }
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, decltype(nullptr), gcq::Combine* cb) const {
    return deq(tid, o, (gcq::gwf::NoProbe*)nullptr, cb);
  }
};

struct LfqQ {
  gcq::lfq::Mpmc m;
  GCQ_NO_LIFETIME
  template <class P = gcq::lfq::NoProbe>
  __device__ __forceinline__ bool enq(uint32_t, uint64_t v, P* p = nullptr, gcq::Combine* ct = nullptr) const {
    return gcq::lfq::enqueue_ptr<P>(m, v, p, ct);
  }
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, decltype(nullptr), gcq::Combine* ct) const {
    return enq(tid, v, (gcq::lfq::NoProbe*)nullptr, ct);
  }
  template <class P = gcq::lfq::NoProbe>
  __device__ __forceinline__ bool deq(uint32_t, uint64_t* o, P* p = nullptr, gcq::Combine* cb = nullptr) const {
    return gcq::lfq::dequeue_ptr<P>(m, o, p, cb);
  }
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, decltype(nullptr), gcq::Combine* cb) const {
    return deq(tid, o, (gcq::lfq::NoProbe*)nullptr, cb);
  }
};

// SFQ, as the reference's harness adapters use it
// (throughout_bench.cpp:251-285):
//  - blocking = 1 (the balanced harness): sfq_enqueue_blocking /
//    sfq_dequeue_blocking, each waiting for its ticket's slot; the wait is
//    bounded (kBlockingSpins checks) and answers BUSY past it, where the
//    reference waits forever;
//  - otherwise enqueue is one non-blocking attempt (BUSY = a legal failed
//    enqueue) and dequeue one non-blocking attempt (the split harness counts
//    BUSY as empty); with exact_empty set the dequeue retries BUSY until it
//    succeeds or observes head >= tail, so an EMPTY result means the queue
//    was empty.
// A Combine passed to a non-blocking operation shares its claim with the
// workgroup's other waves (nb_claim_group).
struct SfqQ {
  static constexpr uint64_t kBlockingSpins = 1ull << 16;
  gcq::sfq::Queue* q;
  uint32_t exact_empty;
  uint32_t blocking;
  GCQ_NO_LIFETIME
  template <class P = gcq::sfq::NoProbe>
  __device__ __forceinline__ bool enq(uint32_t, uint64_t v, P* p = nullptr, gcq::Combine* ct = nullptr) const {
    if (blocking) return gcq::sfq::enqueue_blocking(q, v, kBlockingSpins, p) == gcq::sfq::kSuccess;
    return gcq::sfq::enqueue_nb(q, v, p, ct) == gcq::sfq::kSuccess;
  }
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, decltype(nullptr), gcq::Combine* ct) const {
    return enq(tid, v, (gcq::sfq::NoProbe*)nullptr, ct);
  }
  template <class P = gcq::sfq::NoProbe>
  __device__ __forceinline__ bool deq(uint32_t, uint64_t* o, P* p = nullptr, gcq::Combine* cb = nullptr) const {
    if (blocking) return gcq::sfq::dequeue_blocking(q, o, kBlockingSpins, p) == gcq::sfq::kSuccess;
    if (exact_empty) return gcq::sfq::dequeue_or_empty(q, o, 1u << 20, p, cb);
    return gcq::sfq::dequeue_nb(q, o, p, cb) == gcq::sfq::kSuccess;
  }
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, decltype(nullptr), gcq::Combine* cb) const {
    return deq(tid, o, (gcq::sfq::NoProbe*)nullptr, cb);
  }
};

struct YmcQ {
  gcq::ymc::Queue* q;
  gcq::ymc::Handle* h;
  __device__ __forceinline__ void begin(uint32_t) const { gcq::ymc::begin(q); }
  __device__ __forceinline__ void end(uint32_t tid) const { gcq::ymc::end(q, h, tid); }
  template <class P = gcq::ymc::NoProbe>
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, P* p = nullptr, gcq::Combine* ct = nullptr) const {
    return gcq::ymc::enqueue(q, h, tid, v, p, ct);
  }
  __device__ __forceinline__ bool enq(uint32_t tid, uint64_t v, decltype(nullptr), gcq::Combine* ct) const {
    return enq(tid, v, (gcq::ymc::NoProbe*)nullptr, ct);
  }
  template <class P = gcq::ymc::NoProbe>
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, P* p = nullptr, gcq::Combine* = nullptr) const {
    const uint64_t v = gcq::ymc::dequeue(q, h, tid, p);
    if (v == gcq::ymc::kEmpty) return false;
    *o = v;
    return true;
  }
  __device__ __forceinline__ bool deq(uint32_t tid, uint64_t* o, decltype(nullptr), gcq::Combine* cb) const {
    return deq(tid, o, (gcq::ymc::NoProbe*)nullptr, cb);
  }
};

}  // namespace gcqk

// Call f(adapter) with the adapter matching the queue kind.
// exact_empty selects SFQ's checked-empty dequeue (correctness harnesses).
// blocking selects SFQ's blocking operations (the balanced throughput harness).
template <class F>
int gcq_dispatch(gcq_queue* q, F&& f, bool exact_empty = true, bool blocking = false) {
  switch (q->cfg.kind) {
    case GCQ_GWFQ: {
      gcqk::GwfQ a{gcq::gwf::mpmc_view(q->dev, q->cfg.capacity, q->cfg.max_threads)};
      return f(a);
    }
    case GCQ_GLFQ: {
      gcqk::LfqQ a{gcq::lfq::mpmc_view(q->dev, q->cfg.capacity)};
      return f(a);
    }
    case GCQ_SFQ: {
      gcqk::SfqQ a{(gcq::sfq::Queue*)q->dev, exact_empty ? 1u : 0u, blocking ? 1u : 0u};
      return f(a);
    }
    case GCQ_WFQ: {
      gcqk::YmcQ a{(gcq::ymc::Queue*)q->dev, (gcq::ymc::Handle*)q->aux};
      return f(a);
    }
  }
  gcqh::set_error("unknown queue kind %d", q->cfg.kind);
  return GCQ_EINVAL;
}
