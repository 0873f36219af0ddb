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
  uint32_t wfq_segment_cells; /* WFQ: cells in the pool, a power of two >= 4096; the pool is a ring of
                                 segments reclaimed while the queue runs, and an enqueue is refused (full)
                                 when the ring has no room (0 = max(capacity, 64 per thread) rounded
                                 up to a power of two) */
  uint32_t poll_backoff;  /* G-WFQ/G-LFQ consumer polite-polling base, s_sleep units (0 = off) */
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
  uint32_t poll_percent;  /* waves with (wave % 100) < poll_percent only load, as idle pollers */
  uint32_t poll_other_line; /* 1 = pollers load a different line than the counters */
} gcq_faa_config;
typedef struct gcq_faa_result {
  double tickets_per_s;   /* counter increments per second (all lanes) */
  double atomics_per_s;   /* device atomic instructions per second */
  double kernel_ms;       /* mean launch duration */
  double loads_per_s;     /* pollers' load instructions per second (one per wave) */
} gcq_faa_result;
/* Atomic roofline microbenchmark; no reference counterpart (the reference
 * measures no peak).  Denominator of the split-role roofline. */
int gcq_faa_peak(const gcq_faa_config* cfg, gcq_faa_result* res);

/* Cost of the gfx950 fence forms used by the queues: mode 0 relaxed store,
 * 1 store + release fence, 2 store + acquire fence, 3 store + seq_cst fence. */
int gcq_fence_cost(uint32_t threads, uint32_t block, uint32_t mode, uint32_t iters, double* ns_per_iter);

/* Latency of one dependent device-scope access: one lane follows a random
 * cycle through 65,536 lines, each address from the previous result.  mode
 * 0: device-scope load, 1: device-scope atomic (fetch-add of 0), 2: load
 * through the caches.  The latency bound of level-synchronous work (a BFS
 * level is a chain of such accesses).  No reference counterpart. */
int gcq_round_trip(uint32_t mode, uint32_t steps, double* ns_per_step);

/* Streaming copy of `bytes` (read + write counted) averaged over iters
 * launches: the measured HBM roofline, and a kernel of known traffic for
 * calibrating HBM byte counters.  No reference counterpart. */
int gcq_hbm_copy(uint64_t bytes, uint32_t iters, double* gbps, double* kernel_ms);

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
  uint32_t solo_claims;       /* 1 = no workgroup combining (G-WFQ, G-LFQ);
                                 0 (default) = idle consumers' claims combined
                                 in grids of 8,192 threads or more, producers'
                                 Tail tickets in grids of 32,768 or more */
} gcq_tb_config;
typedef struct gcq_tb_result {
  uint64_t enq_success, enq_fail, deq_success, deq_empty;
  uint64_t launches;
  double elapsed_ms;        /* event time over the timed window (ref semantics) */
  double kernel_ms_mean;    /* mean duration of one chunk launch */
  double succ_mops;         /* (enq_success + deq_success) / elapsed */
  uint64_t pre_enq_success; /* successes in the untimed warmup and probe launches */
  uint64_t pre_deq_success;
} gcq_tb_result;
/* Timed balanced / split run.  Replaces run_balanced_timed /
 * run_split_timed + balanced_chunk_kernel / split_chunk_kernel
 * (throughout_bench.cpp:810-882, 959-1051). */
int gcq_tb_run(gcq_queue* q, const gcq_tb_config* cfg, gcq_tb_result* res);

/* Dequeue everything a quiescent queue holds (no operation in flight) and
 * return the item count in *items, so that a caller can check conservation
 * after a timed run: enqueued = dequeued + drained.  No reference
 * counterpart (the reference's harnesses never account for leftovers,
 * throughout_bench.cpp:959-1051). */
int gcq_queue_drain(gcq_queue* q, uint32_t threads, uint64_t* items);

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

/* ------------------------------------------------------------------ graphs */
/* CSR int32 graph resident in HBM (a host copy is kept for inspection). */
typedef struct gcq_graph gcq_graph;
/* R x C lattice, each lattice edge kept with probability `keep`, plus a short
 * diagonal per vertex with probability `diag`; the reference builds the full
 * lattice only (build_grid, bfs_bench.cpp:245-256). */
int gcq_graph_lattice(int32_t rows, int32_t cols, double keep, double diag, uint64_t seed, gcq_graph** out);
/* R-MAT, 2^scale vertices, edge_factor * 2^scale generated edges,
 * symmetrised and deduplicated.  No reference counterpart. */
int gcq_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                   gcq_graph** out);
/* Load a graph file with the reference loaders' rules (load_mtx / load_snap,
 * bfs_bench.cpp:202-243; format chosen by the ".mtx" suffix under
 * GCQ_GRAPH_AUTO as at bfs_bench.cpp:691-692).  Matrix Market: banner,
 * "rows cols nnz", then nnz 1-based "i j [value]" entries; self loops and
 * out-of-range entries are skipped; V = max(rows, cols).  SNAP: "u v" lines
 * with '#'/'%' comments; self loops skipped; ids shifted down by one when
 * the smallest id is 1; V = largest id + 1.  `undirected` stores both
 * directions.  The CSR is sorted and deduplicated.  Unlike the reference, a
 * missing banner or size line, an empty graph, or more than 2^31 - 2
 * vertices is an error (GCQ_EINVAL) rather than an empty graph. */
enum gcq_graph_format { GCQ_GRAPH_AUTO = 0, GCQ_GRAPH_MTX = 1, GCQ_GRAPH_SNAP = 2 };
int gcq_graph_load(const char* path, int32_t format, int32_t undirected, gcq_graph** out);
/* Upload a caller CSR (sorted, symmetric) to HBM; replaces the CSR H2D copy
 * of bfs_bench.cpp:703-708. */
int gcq_graph_from_csr(int32_t nv, int64_t ne, const int32_t* offsets, const int32_t* indices, gcq_graph** out);
int gcq_graph_info(const gcq_graph* g, int32_t* nv, int64_t* ne);
int gcq_graph_copy_csr(const gcq_graph* g, int32_t* offsets, int32_t* indices);
int32_t gcq_graph_pick_source(const gcq_graph* g, uint64_t seed, uint64_t k);
int gcq_graph_destroy(gcq_graph* g);

/* ------------------------------------------------------------------ BFS */
enum gcq_bfs_mode {
  GCQ_BFS_HOST_LOOP = 0,  /* one launch + host sync per level (bfs_bench.cpp:583-623) */
  GCQ_BFS_PERSISTENT = 1, /* one launch, device-side level barrier */
  GCQ_BFS_ASYNC = 2,      /* one launch, no level barrier: one frontier queue, a vertex whose depth falls
                             (unsigned atomic min) is pushed again, the run ends when no pushed item is left
                             unexpanded.  Same depths bit for bit; edges_scanned counts every expansion,
                             repeated ones included.  Ring frontiers only (G-WFQ, G-LFQ, SFQ). */
  GCQ_BFS_BITMAP = 3      /* as PERSISTENT, but a vertex is claimed by fetch-or in a visited bitmap (one
                             bit per vertex, probed through the caches first) and its depth written by a
                             plain store, instead of a CAS on its depth word: visited neighbours cost a
                             cached load rather than an atomic.  Same depths bit for bit. */
};
typedef struct gcq_bfs_config {
  int32_t kind;       /* frontier queue: enum gcq_kind */
  uint32_t threads;   /* worker threads, multiple of 64 (ref --threads 8192) */
  uint32_t block;     /* (ref --block 256) */
  int32_t mode;       /* enum gcq_bfs_mode */
} gcq_bfs_config;
typedef struct gcq_bfs_result {
  uint64_t edges_scanned;  /* sum of degrees of dequeued vertices (ref edges_total) */
  uint64_t reached;
  int32_t levels;
  int32_t pad;
  double kernel_ms;        /* device time of the traversal (ref kernel_ms) */
  double total_ms;         /* including queue/level reset (ref total_ms) */
} gcq_bfs_result;
/* One BFS from `src` (out of range -> 0, bfs_bench.cpp:555).  depth_out, if
 * not NULL, receives nv int32 depths (-1 = unreached).  Replaces run_bfs +
 * bfs_expand_kernel + bfs_hub_kernel (bfs_bench.cpp:327-633). */
int gcq_bfs_run(gcq_graph* g, int32_t src, const gcq_bfs_config* cfg, int32_t* depth_out, gcq_bfs_result* res);

/* Allocation-free BFS.  No counterpart in the reference: this entry point is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * gcq_bfs_run's persistent kernel with G-LFQ frontiers over a CSR the caller
 * holds in device memory (d_offsets[nv + 1], d_indices[ne], sorted,
 * symmetric) and writes nv int32 depths (-1 = unreached) to the caller's
 * device buffer d_depth.  It calls no hipMalloc/hipFree: its frontier rings
 * and control block live in a fixed device array of
 * GCQ_BFS_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends.  Limits: 1 <= nv <= 2^21 vertices and threads <= 2^21
 * (GCQ_EINVAL beyond).  stream is a hipStream_t (NULL = the default stream);
 * the call waits for the traversal and fills res (kernel_ms is host wall time
 * around the launch sequence).  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU race; the array keeps the last
 * call's contents and each call re-initialises what it uses. */
#define GCQ_BFS_NOALLOC_SCRATCH_BYTES 67113984ull
int gcq_bfs_run_noalloc(const int32_t* d_offsets, const int32_t* d_indices, int32_t nv, int64_t ne, int32_t src,
                        uint32_t threads, uint32_t block, void* stream, int32_t* d_depth, gcq_bfs_result* res);

/* ------------------------------------------------------------------ batched BFS
 * No counterpart in the reference (its BFS harness runs one query per
 * process, bfs_bench.cpp:653-800).  Many traversals, over any of a set of
 * resident graphs, advance together in one persistent kernel that pulls
 * (query slot, vertex) items through one shared frontier queue; each
 * query's depth array is bit-exact with a plain run. */
typedef struct gcq_batch gcq_batch;
typedef struct gcq_batch_config {
  int32_t kind;           /* frontier queue: enum gcq_kind */
  uint32_t threads;       /* persistent grid, multiple of block; all blocks co-resident */
  uint32_t block;         /* 64..1024, multiple of 64 */
  uint32_t pop_width;     /* aggregation width: lanes of a wave that dequeue per step (8..64) */
  uint32_t flush_at;      /* staged discoveries per wave before an aggregated enqueue (8..64) */
  uint32_t capacity_log2; /* frontier ring size; 0 = smallest power of two >= sum of query V */
  uint32_t use_graph;     /* capture the launch sequence in a hipGraph and replay it */
  uint32_t multi_source;  /* 1: queries on one graph run in groups of up to 64 as one bit-parallel traversal
                             (a (group, vertex) item per vertex and level, with a 64-bit mask of the member
                             queries that reach it); same depths bit for bit.  0: one item per (query, vertex) */
} gcq_batch_config;
typedef struct gcq_query_result {
  uint64_t reached;
  uint64_t edges_scanned;
  int32_t levels;
  int32_t status;         /* 0 = done */
} gcq_query_result;
typedef struct gcq_batch_result {
  double total_ms;        /* host wall time of the whole batch, resets and statistics included */
  double kernel_ms;       /* device time of the traversal kernel */
  uint64_t edges_scanned; /* over all queries */
  int32_t levels;         /* global levels the batch ran */
  int32_t graph_replayed; /* 1 if the launch was a replay of a captured graph */
} gcq_batch_result;
/* A batch context over `ngraphs` graphs (borrowed: they must outlive it). */
int gcq_batch_create(gcq_graph* const* graphs, int32_t ngraphs, gcq_batch** out);
void gcq_batch_config_default(gcq_batch_config* cfg);
/* Run nq queries (graph_ids[i], sources[i]) as one batch.  depth_out, if
 * not NULL, holds nq host pointers (each NULL or room for that graph's nv
 * int32 depths); per_query, if not NULL, receives nq results. */
int gcq_batch_run(gcq_batch* b, const gcq_batch_config* cfg, int32_t nq, const int32_t* graph_ids,
                  const int32_t* sources, int32_t* const* depth_out, gcq_query_result* per_query,
                  gcq_batch_result* res);
/* Project-written autotuner: times the batch under candidate configurations
 * (block, threads, pop_width, flush_at, capacity) and writes the fastest to
 * best; trials_out, if not NULL, receives up to max_trials (config, ms) rows. */
typedef struct gcq_tune_trial {
  gcq_batch_config cfg;
  double ms;
} gcq_tune_trial;
int gcq_batch_autotune(gcq_batch* b, int32_t kind, int32_t nq, const int32_t* graph_ids, const int32_t* sources,
                       int32_t reps, gcq_batch_config* best, gcq_tune_trial* trials_out, int32_t max_trials,
                       int32_t* ntrials);
int gcq_batch_destroy(gcq_batch* b);

/* ------------------------------------------------------------------ ray tracing
 * Queue-compacted Whitted tracer of rt_final.cpp (types :343-362, shading
 * :517-645, scenes :650-684, camera :849-859): spheres and planes, one
 * directional light, mirror bounces up to max_depth, sky on a miss.  Each
 * tile has one queue holding the local pixel indices of its live bounce
 * rays; a pixel's contributions are summed in bounce order, so the image is
 * independent of scheduling.  Replaces k_generate + k_trace + main. */
typedef struct gcq_rt_sphere { float c[3]; float r; float albedo[3]; float refl; } gcq_rt_sphere;  /* rt_final.cpp:343 */
typedef struct gcq_rt_plane { float n[3]; float d; float albedo[3]; float refl; } gcq_rt_plane;    /* rt_final.cpp:344 */
typedef struct gcq_rt_scene_desc {
  const gcq_rt_sphere* spheres;
  int32_t nspheres;
  int32_t nplanes;
  const gcq_rt_plane* planes;
  float cam_o[3];
  float cam_f[3], cam_r[3], cam_u[3]; /* forward/right/up as given; normalised on the host (rt_final.cpp:856-859) */
  float light[3];                     /* direction towards the light, normalised in the kernel (:560) */
  float ambient[3];                   /* (:561) */
} gcq_rt_scene_desc;
typedef struct gcq_rt_scene gcq_rt_scene;
enum gcq_rt_preset {
  GCQ_RT_COMPLEX = 0,  /* build_scene_complex + its camera (rt_final.cpp:650-671, :850-852) */
  GCQ_RT_CORNELL = 1,  /* build_scene_cornell + its camera (:673-684, :853-855) */
  GCQ_RT_RANDOM = 2    /* n random spheres on a ground plane (no counterpart in the reference) */
};
int gcq_rt_scene_create(const gcq_rt_scene_desc* desc, gcq_rt_scene** out);
/* nspheres and seed are used by GCQ_RT_RANDOM only. */
int gcq_rt_scene_preset(int32_t preset, int32_t nspheres, uint64_t seed, gcq_rt_scene** out);
/* Fills desc with pointers into the scene's host copy (valid until destroy). */
int gcq_rt_scene_get(const gcq_rt_scene* s, gcq_rt_scene_desc* desc);
int gcq_rt_scene_destroy(gcq_rt_scene* s);

typedef struct gcq_rt_config {
  int32_t kind;              /* per-tile queue: enum gcq_kind */
  uint32_t threads_per_tile; /* workers per tile, multiple of 64 (ref argv[3], default 1024) */
  uint32_t block;            /* (ref TPB 256) */
  int32_t max_depth;         /* bounces (ref argv[2], default 4) */
  int32_t width, height;     /* full image (camera aspect and pixel centres) */
  int32_t tile_w, tile_h;    /* tiling of the rendered region (ref 160 x 90) */
  int32_t x0, y0, w, h;      /* rendered region; w = h = 0 means the whole image */
} gcq_rt_config;
typedef struct gcq_rt_result {
  uint64_t primary_rays;
  uint64_t total_rays;       /* rays traced, bounces included (ref traced) */
  double kernel_ms;          /* the one trace launch */
  double total_ms;           /* host wall time, resets and read-back included */
  int32_t tiles;
  int32_t pad;
} gcq_rt_result;
void gcq_rt_config_default(gcq_rt_config* cfg);
/* Render cfg's region into img_out (w * h RGB float triples, row-major). */
int gcq_rt_render(gcq_rt_scene* s, const gcq_rt_config* cfg, float* img_out, gcq_rt_result* res);
/* The stream-compaction comparator (concurrent_rt/baselines/rt_compaction.cpp:
 * 219-472, 748-833): one launch per bounce traces every live ray and packs
 * the continuing ones for the next bounce with a prefix sum, with the same
 * shading as gcq_rt_render, so the image is the same.  Uses cfg's image,
 * region, max_depth and block; kind, threads_per_tile and the tiling are
 * ignored.  kernel_ms is the frame: first launch to last kernel end,
 * including the host's read of each bounce's ray count. */
int gcq_rt_render_compact(gcq_rt_scene* s, const gcq_rt_config* cfg, float* img_out, gcq_rt_result* res);
/* Write a w x h RGB float image as an 8-bit binary PPM, each channel
 * truncated from min(255, max(0, x * 255)) as save_ppm does
 * (rt_final.cpp:689-700). */
int gcq_rt_write_ppm(const float* img, int32_t w, int32_t h, const char* path);
/* Append one result row to a CSV file, writing the header first when the
 * file does not exist.  GCQ_RT_CSV_QUEUE: the columns of rt_queue_perf.csv
 * (rt_final.cpp:972-1001; every header column filled, GEN_MS = 0 because
 * generation runs inside the trace launch).  GCQ_RT_CSV_COMPACT: the columns
 * of rt_performance.csv (baselines/rt_compaction.cpp:849-864; THREADS is
 * the block size). */
enum gcq_rt_csv_style { GCQ_RT_CSV_QUEUE = 0, GCQ_RT_CSV_COMPACT = 1 };
typedef struct gcq_rt_csv_row {
  const char* gpu;         /* device name */
  const char* queue;       /* "GWFQ", "GLFQ", "WFQ", "SFQ" or "COMPACT" */
  const char* scene;       /* "complex", "cornell", ... (queue style only) */
  uint32_t threads;        /* threads per tile (queue style) or block (compaction style) */
  int32_t bounces;         /* max_depth (queue style only) */
  const gcq_rt_result* res;
} gcq_rt_csv_row;
int gcq_rt_append_csv(const char* path, int32_t style, const gcq_rt_csv_row* row);

/* ------------------------------------------------------------------ serving
 * No counterpart in the reference.  A continuously batched server: BFS
 * queries on resident graphs and RT tile jobs on a resident scene are
 * admitted by a host loop as slots and device memory free up and run in one
 * persistent kernel (a BFS sub-grid advancing every admitted query level by
 * level, an RT sub-grid tracing admitted tiles).  Requests reach the kernel
 * through two device-resident submission queues (G-LFQ rings in HBM, one
 * for BFS queries and one for tile jobs), fed by one gate wave of the kernel
 * from a doorbell ring in pinned host memory; the BFS sub-grid takes queries
 * at its level barriers, RT waves take tile jobs as soon as they look for
 * work.  Completions come back through a pinned completion ring.  While the
 * server runs nothing else is launched, so no stream can queue behind the
 * persistent kernel. */
typedef struct gcq_server gcq_server;
enum gcq_req_type { GCQ_REQ_BFS = 0, GCQ_REQ_RT = 1 };
typedef struct gcq_request {
  int32_t type;           /* enum gcq_req_type */
  int32_t graph;          /* BFS: index into the server's graphs */
  int32_t src;            /* BFS: source vertex (out of range -> 0) */
  int32_t x0, y0, w, h;   /* RT: tile of the server's frame */
  int32_t keep;           /* 1: keep the result (in keep_bytes) for gcq_server_fetch after stop */
  uint64_t id;            /* caller's id, echoed in the completion */
  double arrival_ms;      /* trace runs: arrival time after the trace start */
} gcq_request;
typedef struct gcq_completion {
  uint64_t id;
  int32_t type, status;   /* status 0 = done */
  int32_t levels;         /* BFS: levels of the traversal */
  int32_t pad;
  uint64_t rays;          /* RT: rays traced */
  double arrival_ms, admit_ms, done_ms;  /* host clock, after the trace start */
} gcq_completion;
typedef struct gcq_server_config {
  int32_t kind;            /* BFS frontier queue (enum gcq_kind) */
  uint32_t bfs_threads;    /* BFS sub-grid, multiple of block */
  uint32_t rt_threads;     /* RT sub-grid, multiple of block; its first wave is the gate (>= 128 with a
                              scene; 0 without one adds a block for the gate) */
  uint32_t block;
  uint32_t bfs_slots;      /* concurrent BFS queries */
  uint32_t rt_slots;       /* concurrent RT tile jobs */
  uint32_t frontier_log2;  /* frontier ring size; in-flight queries' V sum stays below it */
  int32_t max_depth;       /* RT bounces */
  uint32_t pop_width, flush_at;
  uint64_t budget_bytes;   /* device memory for request state: depth arrays, ray records, images */
  uint64_t keep_bytes;     /* separate device memory for results kept for gcq_server_fetch */
  uint32_t max_requests;   /* submission / completion ring entries for the server's lifetime */
  uint32_t small_log2;     /* small-graph lane: queries on graphs of at most 2^small_log2 vertices run in a
                              second BFS sub-grid with its own level barrier (0 = one lane for all) */
  uint32_t small_threads;  /* the second lane's threads, taken out of bfs_threads (multiple of block) */
  uint32_t multi_source;   /* 1: queries on one graph share bit-parallel groups of 64 (ceil(bfs_slots / 64)
                              groups per graph, 24 B per vertex each, outside budget_bytes); each query's
                              depths stay bit-exact.  0: one (query, vertex) item per discovery */
} gcq_server_config;
void gcq_server_config_default(gcq_server_config* cfg);
/* graphs and scene are borrowed and must outlive the server; scene may be
 * NULL when no RT request will be sent. */
int gcq_server_create(gcq_graph* const* graphs, int32_t ngraphs, gcq_rt_scene* scene, int32_t frame_w,
                      int32_t frame_h, const gcq_server_config* cfg, gcq_server** out);
int gcq_server_start(gcq_server* s);
/* Queue a request on the host; it is admitted as soon as a slot and memory allow. */
int gcq_server_submit(gcq_server* s, const gcq_request* r);
/* Admit what fits and collect up to max completions. */
int gcq_server_poll(gcq_server* s, gcq_completion* out, int32_t max, int32_t* n);
/* Wait for every submitted request, then end the persistent kernel. */
int gcq_server_stop(gcq_server* s);
/* After stop: the depth array (BFS, nv int32) or image (RT, w*h RGB floats) of a kept request. */
int gcq_server_fetch(gcq_server* s, uint64_t id, void* out);
/* Start, submit reqs[i] at reqs[i].arrival_ms, poll until all are done, stop;
 * out[i] is the completion of reqs[i]. */
int gcq_server_run_trace(gcq_server* s, const gcq_request* reqs, int32_t n, gcq_completion* out);
int gcq_server_destroy(gcq_server* s);
/* The plain comparator: the same trace served one request at a time with
 * one launch (gcq_bfs_run / gcq_rt_render) per request. */
int gcq_plain_run_trace(gcq_graph* const* graphs, int32_t ngraphs, gcq_rt_scene* scene, int32_t frame_w,
                        int32_t frame_h, const gcq_bfs_config* bcfg, const gcq_rt_config* rcfg,
                        const gcq_request* reqs, int32_t n, gcq_completion* out);

#ifdef __cplusplus
}
#endif
#endif /* GCQ_H */
