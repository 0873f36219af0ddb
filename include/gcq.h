/* gcq.h — C ABI of libgcq, the MI355X-native implementation of gCQ's GPU
 * MPMC FIFO queues (G-WFQ, G-LFQ, WFQ-YMC, SFQ) and of the queue-driven
 * workloads built on them.
 *
 * The reference (gosLp/gCQ) exposes no host ABI: each harness is a
 * monolithic main() that calls the header-only device API.  Every entry point
 * below names the reference host code it replaces (file:line).  Conventions:
 *   - return 0 on success, a negative GCQ_E* code on failure; nothing calls
 *     exit() and no C++ exception crosses the ABI (gcq_strerror / gcq_last_error
 *     give the message);
 *   - plain pointers and sizes only; "host" buffers are ordinary host memory,
 *     "device" pointers come from the library;
 *   - a handle is not thread-safe; each handle owns one HIP stream.
 */
#ifndef GCQ_H
#define GCQ_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define GCQ_OK 0
#define GCQ_EINVAL -1     /* bad argument */
#define GCQ_ENOMEM -2     /* device allocation failed */
#define GCQ_EHIP -3       /* HIP runtime error */
#define GCQ_ETIMEOUT -4   /* watchdog / spin limit hit */
#define GCQ_EFULL -5      /* capacity exhausted */
#define GCQ_EVERIFY -6    /* a device-side self-check failed */
#define GCQ_ENODEV -7     /* no usable GPU */

const char* gcq_strerror(int rc);
const char* gcq_last_error(void);
const char* gcq_version(void);
/* Device name, CU count and HBM size of device 0 (reference prints these at
 * throughout_bench.cpp:1058-1062, bfs_bench.cpp:677-682). */
int gcq_device_info(char* name, size_t name_len, int* cu_count, uint64_t* hbm_bytes);

/* ------------------------------------------------------------------ queues */
enum gcq_kind { GCQ_GWFQ = 0, GCQ_GLFQ = 1, GCQ_WFQ = 2, GCQ_SFQ = 3 };

typedef struct gcq_queue gcq_queue;

typedef struct gcq_queue_config {
  int32_t kind;           /* enum gcq_kind */
  uint32_t capacity;      /* ring capacity, power of two (G-WFQ/G-LFQ/SFQ) */
  uint32_t max_threads;   /* highest thread id + 1 that may use the queue */
  uint32_t remap;         /* 1 = spread consecutive tickets over 128-B lines */
  uint32_t help_delay;    /* G-WFQ: ops between help turns (ref WF_HELP_DELAY 64) */
  uint32_t enq_patience;  /* G-WFQ fast-path attempts + 1 (ref WF_ENQ_PATIENCE 128) */
  uint32_t deq_patience;  /* (ref WF_DEQ_PATIENCE 256) */
  uint32_t wfq_segment_cells; /* WFQ: cells reserved in the segment pool (0 = auto) */
} gcq_queue_config;

/* Fill a config with the reference defaults for `kind`. */
void gcq_queue_config_default(int kind, gcq_queue_config* cfg);

/* Allocate and initialise a queue in HBM with a parallel init kernel.
 * Replaces create_queue (throughout_bench.cpp:420-442, history.cpp:151-166),
 * wf_queue_host_init (wfqueue_hip_opt.hpp:594-644, 762-764) and
 * sfq_queue_host_init (sfqueue_hip.cpp:356-390). */
int gcq_queue_create(const gcq_queue_config* cfg, gcq_queue** out);
/* Return the queue to its initial empty state (ref: re-running
 * queue_init_kernel, bfs_bench.cpp:548-551; wf_queue_reset_for_bfs,
 * wfqueue_hip_opt.hpp:665-692). */
int gcq_queue_reset(gcq_queue* q);
/* ref destroy_queue (throughout_bench.cpp:444-453), sfq_queue_destroy
 * (sfqueue_hip.cpp:392-399). */
int gcq_queue_destroy(gcq_queue* q);
/* Device address of the queue object (for callers that launch their own
 * kernels against the device headers) and its size in bytes. */
void* gcq_queue_device_ptr(gcq_queue* q);
uint64_t gcq_queue_bytes(const gcq_queue* q);

/* ------------------------------------------------------------------ FAA peak */
typedef struct gcq_faa_config {
  uint32_t threads, block;
  uint32_t counters;      /* K distinct counters (1 = single-counter peak) */
  uint32_t stride_bytes;  /* distance between counters */
  uint32_t aggregate;     /* 1 = one FAA per wave (ballot + prefix), 0 = per lane */
  uint32_t iters;         /* FAAs per thread per launch */
  uint32_t launches;
} gcq_faa_config;
typedef struct gcq_faa_result {
  double tickets_per_s;   /* counter increments per second (all lanes) */
  double atomics_per_s;   /* device atomic instructions per second */
  double kernel_ms;       /* mean launch duration */
} gcq_faa_result;
/* Atomic roofline microbenchmark; no reference counterpart (the reference
 * measures no peak).  Denominator of the split-role roofline. */
int gcq_faa_peak(const gcq_faa_config* cfg, gcq_faa_result* res);

/* Cost of the gfx950 fence forms used by the queues: mode 0 relaxed store,
 * 1 store + release fence, 2 store + acquire fence, 3 store + seq_cst fence. */
int gcq_fence_cost(uint32_t threads, uint32_t block, uint32_t mode, uint32_t iters, double* ns_per_iter);

/* ------------------------------------------------------------------ throughput */
enum gcq_tb_mode { GCQ_TB_BALANCED = 0, GCQ_TB_SPLIT = 1 };
typedef struct gcq_tb_config {
  int32_t mode;             /* enum gcq_tb_mode */
  uint32_t threads, block;
  uint32_t chunk_ops;       /* ops per thread per launch (ref CHUNK_OPS) */
  uint32_t producer_percent;/* split: producer iff tid*100/threads < p */
  uint32_t warmup_ms, run_ms;
  uint32_t launches_per_sync; /* async launches between host time checks (0 = auto) */
  uint32_t warmup_launches;   /* > 0 with timed_launches: launch-count mode instead */
  uint32_t timed_launches;    /* exactly this many timed launches (0 = time mode) */
} gcq_tb_config;
typedef struct gcq_tb_result {
  uint64_t enq_success, enq_fail, deq_success, deq_empty;
  uint64_t launches;
  double elapsed_ms;        /* event time over the timed window (ref semantics) */
  double kernel_ms_mean;    /* mean duration of one chunk launch */
  double succ_mops;         /* (enq_success + deq_success) / elapsed */
} gcq_tb_result;
/* Timed balanced / split run.  Replaces run_balanced_timed /
 * run_split_timed + balanced_chunk_kernel / split_chunk_kernel
 * (throughout_bench.cpp:810-882, 959-1051). */
int gcq_tb_run(gcq_queue* q, const gcq_tb_config* cfg, gcq_tb_result* res);

/* ------------------------------------------------------------------ FIFO check */
typedef struct gcq_fifo_result {
  uint64_t in_count, out_count;
  uint32_t err_oob, err_dup, err_order, err_missing, err_gt1;
  int32_t verified;
} gcq_fifo_result;
/* Produce-all then drain with exactly-once and per-producer order checks;
 * concurrent = 1 runs producers and consumers in one launch instead.
 * Replaces run_fifo_check_once (throughout_bench.cpp:531-751). */
int gcq_fifo_check(gcq_queue* q, uint32_t threads, uint32_t ops, uint32_t block, int32_t concurrent,
                   gcq_fifo_result* res);

/* ------------------------------------------------------------------ history */
/* One logged operation (reference WFLogEntry, correctness/wf_linlog.hpp:7-14,
 * plus the per-op step count). */
typedef struct gcq_hist_entry {
  uint32_t proc, op;        /* op: 0 = enqueue, 1 = dequeue */
  uint64_t arg, ret;        /* enq: value, ret 1/0; deq: value or 0 = empty */
  uint64_t call, ret_time;  /* logical clock at call / return */
  uint32_t steps;           /* shared-memory steps the operation took */
  uint32_t slow;            /* 1 if it entered the wait-free slow path */
} gcq_hist_entry;
typedef struct gcq_hist_config {
  uint32_t threads, ops, block;
  uint32_t mode;              /* 0 alternating enq/deq, 1 split roles */
  uint32_t producer_percent;
  uint32_t delay_every;       /* adversary: stall threads with tid % delay_every == 0 (0 = off) */
  uint32_t delay_cycles;      /* stall length, between ticket and slot access */
  uint64_t seed;              /* adversary placement salt */
} gcq_hist_config;
/* Run a logged history kernel; entries are written to `out` (host), up to
 * `cap` entries, sorted by (call, ret_time, proc, op).  Replaces
 * alternating_history_kernel / split_history_kernel + wf_log_dump_jsonl
 * (history.cpp:188-283). */
int gcq_hist_run(gcq_queue* q, const gcq_hist_config* cfg, gcq_hist_entry* out, size_t cap, size_t* n_out);
/* Write entries as the reference JSONL schema (history.cpp:271-279). */
int gcq_hist_write_jsonl(const gcq_hist_entry* e, size_t n, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* GCQ_H */
