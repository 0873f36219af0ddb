/* oracle.h — ORACLE (test infrastructure only).  CPU restatements of the
 * reference's algorithms used to check the product (tests/, the smoke check
 * and the bench's cpu_baseline leg).  Nothing in the product links this. */
#ifndef GCQ_ORACLE_H
#define GCQ_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_op {
  uint32_t proc, op; /* op 0 enqueue, 1 dequeue */
  uint64_t arg, ret, call, end;
} orc_op;

#define ORC_OK 1
#define ORC_ILLEGAL 0
#define ORC_UNKNOWN -1
#define ORC_ENGINE_AUTO 0
#define ORC_ENGINE_WGL 1
#define ORC_ENGINE_PATTERN 2

/* FIFO linearizability (model of correctness/lincheck/main.go:90-136). */
int orc_lincheck(const orc_op* ops, size_t n, int engine, double budget_s, char* msg, size_t msg_len);
/* Parse the reference JSONL history schema (history.cpp:271-279); returns the
 * number of records (may exceed cap), -1 open error, -2 parse error. */
long orc_read_jsonl(const char* path, orc_op* out, size_t cap);

/* Host std::atomic restatements of the four queues. */
enum { ORC_GWFQ = 0, ORC_GLFQ = 1, ORC_WFQ = 2, ORC_SFQ = 3 };
typedef struct orc_hist_cfg {
  int32_t kind;
  uint32_t threads, ops, mode, producer_percent;
  uint32_t capacity;
  uint32_t enq_patience, deq_patience; /* G-WFQ / WFQ; 0 = reference default */
  uint32_t ref_slow_rule;              /* G-WFQ: 1 = reference try_enq_slow rule */
  uint32_t delay_every, delay_us;      /* stall threads tid % delay_every == 0 after a ticket */
} orc_hist_cfg;
/* Run a logged history with std::threads; returns the number of entries
 * written (sorted by call time) or a negative value on error. */
long orc_hist_run(const orc_hist_cfg* cfg, orc_op* out, size_t cap, uint32_t* max_steps);

typedef struct orc_tb_cfg {
  int32_t kind;
  uint32_t threads, mode, producer_percent, chunk_ops, warmup_ms, run_ms, capacity;
} orc_tb_cfg;
typedef struct orc_tb_res {
  uint64_t enq_success, enq_fail, deq_success, deq_empty;
  double elapsed_ms, succ_mops;
  uint32_t hw_threads;
} orc_tb_res;
/* Timed balanced/split throughput with std::threads (same split rule). */
int orc_tb_run(const orc_tb_cfg* cfg, orc_tb_res* res);

/* Serial textbook BFS on an int32 CSR; depth -1 = unreached.  Returns the
 * number of reached vertices.  `edges_scanned` = sum of degrees of reached. */
long orc_bfs_serial(int32_t nv, const int32_t* offs, const int32_t* idx, int32_t src, int32_t* depth,
                    uint64_t* edges_scanned);
/* Level-synchronous top-down BFS over `threads` host threads (0 = all). */
long orc_bfs_parallel(int32_t nv, const int32_t* offs, const int32_t* idx, int32_t src, int32_t* depth,
                      int threads);

/* Reference ray tracer image of region (x0, y0, w, h) of a W x H frame:
 * spheres / planes as 8 floats each (centre|normal, radius|d, albedo, refl),
 * cam12 = origin, forward, right, up (not yet normalised), light direction
 * (not yet normalised), ambient.  img gets w*h RGB floats. */
int orc_rt_render(const float* sph, int32_t ns, const float* pl, int32_t np, const float* cam12, const float* light3,
                  const float* ambient3, int32_t W, int32_t H, int32_t x0, int32_t y0, int32_t w, int32_t h,
                  int32_t max_depth, int32_t threads, float* img, uint64_t* rays_traced);

#ifdef __cplusplus
}
#endif
#endif
