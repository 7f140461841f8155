/*
 * bfsx — MI355X (gfx950) breadth-first-search engine, C ABI.
 *
 * This is the drop-in boundary for the reference's GPU path,
 * src/run_bfs_gpu.hpp (GPURun, :534-542, and the kernels it launches).  All
 * entry points take plain pointers and sizes, return a status code
 * (BFSX_OK = 0) and never let a C++ exception cross the ABI; the message of
 * the last failure is available from bfsx_last_error().
 *
 * A context owns one HIP device, one stream (default priority) and all device
 * memory.  The graph stays resident in HBM across runs (the reference
 * re-uploads it every run, run_bfs_gpu.hpp:620-668, outside its timer).  A
 * context is single-threaded: calls on one context must not overlap.
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference repository):
 *   bfsx_create / bfsx_destroy    queue setup and teardown,
 *                                 src/run_bfs_gpu.hpp:546-557, :732-746
 *   bfsx_upload_csc               malloc_device + copyToDevice of col_ptr and
 *                                 row_idx, src/run_bfs_gpu.hpp:620-665
 *   bfsx_run                      one run of the timed level loop
 *                                 (explore + levelgen per level),
 *                                 src/run_bfs_gpu.hpp:671-730
 *   bfsx_run_batch                the num_runs loop of GPURun over roots,
 *                                 src/run_bfs_gpu.hpp:614-747
 *   bfsx_set_params               SM_FACTOR and work-group size,
 *                                 src/CMakeLists.txt:12, run_bfs_gpu.hpp:194-197
 */
#ifndef BFSX_H
#define BFSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BFSX_OK 0
#define BFSX_ERR_ARG 1          /* invalid argument (bad root, size, width) */
#define BFSX_ERR_DEVICE 2       /* HIP runtime error or no usable device */
#define BFSX_ERR_NOMEM 3        /* device or host allocation failed */
#define BFSX_ERR_STATE 4        /* call out of order (e.g. run before upload) */
#define BFSX_ERR_IO 5           /* file missing, empty, short or inconsistent */
#define BFSX_ERR_OVERFLOW 6     /* a result does not fit its output type */
#define BFSX_ERR_UNSUPPORTED 7  /* strategy not valid for this graph */

/* ---- strategies (the reference's compile-time switch,
 *      src/CMakeLists.txt:10-11,63-64, made a run-time enum) -------------- */
#define BFSX_LLB 0      /* local load balance: ExploreNeighbours, :388-477 */
#define BFSX_GLB 1      /* global load balance: ExploreNeighbours1..4, :175-385 */
#define BFSX_SLB 2      /* strided local load balance: :416-417 */
#define BFSX_PULL 3     /* bottom-up step (symmetric graphs only) */
#define BFSX_LEARNED 4  /* per-level choice of the four by the MLP scheduler */
#define BFSX_BEAMER 5   /* direction-optimising rule of Beamer, Asanovic and
                           Patterson (SC'12), alpha = 14, beta = 24: bottom-up
                           once the frontier's edges exceed 1/alpha of the
                           unexplored vertices' edges, top-down again once
                           the frontier shrinks below n/beta; top-down levels
                           run GLB when the frontier has at least n/256
                           edges, else LLB (symmetric graphs; elsewhere
                           top-down only).  The hand-set rule the learned
                           scheduler is compared with. */
#define BFSX_N_STEP_KINDS 4

/* graph flags */
#define BFSX_GRAPH_SYMMETRIC 1u /* every stored edge u->v has v->u */
#define BFSX_GRAPH_HUB_FIRST 2u /* adjacency lists ordered hub first (below) */
#define BFSX_GRAPH_PULL_COPY 4u /* bottom-up steps read a hub-first copy (below) */

/* run flags */
#define BFSX_RUN_GRAPH_LOOP 1u  /* capture the level loop in a hipGraph with a
                                   device-side termination flag */
#define BFSX_RUN_TRACE 2u       /* record per-level statistics and timings */
#define BFSX_RUN_STEP_TIMES 4u  /* record per-level step times only (no
                                   extra per-vertex work in the kernels) */
#define BFSX_RUN_DEVICE_LOOP 8u /* device-driven thin levels: runs of top-down
                                   levels whose frontier edges stay below the
                                   mark threshold (mark_ratio_q10) and whose
                                   kind is LLB or SLB (fixed, or the learned
                                   scheduler's choice, made on the device)
                                   execute in one cooperative launch with
                                   grid-wide barriers between levels; heavy
                                   levels and GLB or PULL steps run as host-
                                   issued steps.  Any strategy but BFSX_BEAMER;
                                   not with traces or parents. */

typedef struct bfsx_ctx bfsx_ctx;

typedef struct bfsx_graph_info {
  uint32_t n_rows;       /* vertices expanded (the reference's numRows) */
  uint32_t n_cols;       /* vertices / level-array length (numCols) */
  uint64_t nnz;          /* stored edges */
  int32_t col_ptr_bytes; /* 4 or 8 */
  uint32_t flags;        /* BFSX_GRAPH_* */
  uint32_t max_degree;
  uint32_t n_isolated;   /* vertices < n_rows with no stored edge */
  uint32_t pull_hubs;    /* top-degree vertices that bottom-up steps probe in
                            a separate hub bitmap (BFSX_ORDER_PULL_HUB_FIRST
                            builds it; 0: none) */
  uint32_t hot_vertices; /* top-degree vertices that heavy top-down LLB/SLB
                            levels probe in a separate visited bitmap
                            (BFSX_ORDER_PULL_HUB_FIRST builds it on graphs of
                            2^24 vertices or more; 0: none) */
} bfsx_graph_info;

typedef struct bfsx_run_stats {
  double ms;             /* level loop only: after upload/initialisation,
                            before the copy of levels to the host
                            (src/run_bfs_gpu.hpp:679,728) */
  uint64_t edges_count;  /* the reference's edgesCount: sum of out-degree
                            over reached vertices < n_rows
                            (src/run_bfs_cpu.hpp:36,84) */
  uint32_t n_levels;     /* level iterations that expanded a frontier */
  uint32_t max_level;    /* deepest level reached */
  uint64_t reached;      /* vertices with a level >= 0 */
  uint32_t steps[BFSX_N_STEP_KINDS]; /* how many levels used each step kind */
} bfsx_run_stats;

/* one level of the last run (BFSX_RUN_TRACE) */
typedef struct bfsx_level_record {
  uint32_t level;           /* frontier level being expanded */
  uint32_t choice;          /* BFSX_LLB..BFSX_PULL */
  uint64_t frontier_vertices;
  uint64_t frontier_edges;  /* sum of out-degree of the frontier */
  uint64_t unvisited_vertices;
  uint64_t unvisited_edges;
  uint64_t discovered;      /* vertices found by this step */
  double us;                /* device time of this step */
  uint64_t pull_scanned;    /* adjacency entries read by a bottom-up step */
  double kernel_us;         /* device time of the step's main kernel
                               (k_llb / k_glb_expand / k_pull) */
  /* learned strategy: the scheduler decision that picked this level's kind
   * (made at the end of the previous level); sched_evaluated = 1 when the
   * network ran for it, 0 when the last decision was kept (similar frontier)
   * or for the root level, which runs LLB without an evaluation */
  uint32_t sched_evaluated;
  float sched_in[8];        /* network inputs of the decision in force */
  float sched_out[4];       /* its outputs: predicted log2 time of LLB, GLB,
                               SLB, PULL relative to LLB */
} bfsx_level_record;

/* launch-shape parameters (the SM_FACTOR / work-group knobs) */
typedef struct bfsx_params {
  uint32_t llb_wg;          /* work-group size for LLB/SLB: 64..1024, pow2 */
  uint32_t glb_wg;          /* work-group size of the GLB expand kernel */
  uint32_t glb_blocks_per_cu; /* persistent GLB grid = CUs * this (SM_FACTOR) */
  uint32_t glb_chunk;       /* edges per GLB work-queue item */
  uint32_t hub_degree;      /* LLB/SLB: frontier vertices with at least this
                               degree are expanded by a whole wave (0 = off) */
  uint32_t pull_wg;         /* work-group size of the bottom-up step */
  uint32_t graph_levels;    /* levels per captured hipGraph launch */
  uint32_t mark_ratio_q10;  /* top-down steps whose frontier edges E (exact
                               out-degree sum, kept on the device) reach
                               n_cols * mark_ratio_q10 / 1024 with at least
                               8 edges per frontier vertex, or reach
                               max(that, n_cols / 16) in any case, mark
                               neighbours with plain byte stores and compact
                               afterwards; the others discover exact-once
                               with atomics (0 = always mark,
                               0xFFFFFFFF = never) */
  uint32_t bucket_ratio_q10; /* a GLB step with at least n_cols/256
                               frontier edges (or that would mark) uses
                               bucketed discovery (sequential writes of
                               every target, no random stores or atomics)
                               when the unvisited vertices' edges are at
                               least nnz * bucket_ratio_q10 / 1024, i.e. when
                               most frontier edges lead to new vertices
                               (0 = always, 0xFFFFFFFF = never; not with
                               mark_ratio_q10 = 0xFFFFFFFF or parents;
                               graphs of up to 2^27 vertices) */
  /* batched searches (bfsx_run_multi) */
  uint32_t ms_relabel;      /* 1: search a copy of the graph with vertices
                               renamed by degree (hubs first; made once per
                               graph, symmetric graphs only), 0: as stored */
  uint32_t ms_push_divisor; /* a level runs top-down while its frontier's
                               edges * this < nnz, else bottom-up (1..1024) */
  uint32_t ms_pull_hub;     /* bottom-up levels give vertices with more
                               neighbours than this a work-group each (32..65536) */
  uint32_t ms_graph_levels; /* levels per captured hipGraph launch (2..64, even) */
  uint32_t coop_groups;     /* BFSX_RUN_DEVICE_LOOP: work-groups of the
                               cooperative launch (1..4096, capped by what
                               fits resident; default 256; grid-wide
                               barriers cost more the more groups take
                               part) */
  uint32_t speculate;       /* learned scheduler, untimed runs: launch the
                               next level's step before the current level's
                               decision is read back (a bottom-up step after
                               a level that wrote the next frontier's bitmap,
                               a thin LLB/SLB step after a top-down level);
                               it runs only if the decision matches, else
                               exits without a write (1, default) or not (0) */
} bfsx_params;

/* ---- context ------------------------------------------------------------ */
int bfsx_create(int device, bfsx_ctx **out);
void bfsx_destroy(bfsx_ctx *ctx);
/* message of the context's last failure; with NULL, that of the calling
 * thread's last bfsx_run_noalloc */
const char *bfsx_last_error(const bfsx_ctx *ctx);
const char *bfsx_status_string(int status);
int bfsx_device_info(bfsx_ctx *ctx, char *name, size_t name_len,
                     int *compute_units, uint64_t *hbm_bytes);

/* ---- graph -------------------------------------------------------------- */
/* Uploads a CSC graph in the layout of src/graph_load.hpp: col_ptr has
 * n_cols+1 entries of col_ptr_bytes (4 or 8) each, row_idx has nnz entries.
 * Vertices >= n_rows are never expanded (src/run_bfs_cpu.hpp:31). */
int bfsx_upload_csc(bfsx_ctx *ctx, uint32_t n_cols, uint32_t n_rows,
                    const void *col_ptr, int col_ptr_bytes,
                    const uint32_t *row_idx, uint64_t nnz, uint32_t flags);
/* Builds an R-MAT graph directly in HBM: edge_factor * 2^scale generated
 * edges with quadrant probabilities a, b, c (d = 1-a-b-c), a seeded vertex
 * permutation, self-loops and duplicates dropped, symmetrized. */
int bfsx_generate_rmat(bfsx_ctx *ctx, int scale, int edge_factor, double a,
                       double b, double c, uint64_t seed_graph,
                       uint64_t seed_perm, int col_ptr_bytes, uint64_t *nnz);
/* Builds a side x side 4-neighbour grid in HBM with each undirected edge
 * dropped with probability drop_prob and n_long random long-range edges. */
int bfsx_generate_grid(bfsx_ctx *ctx, uint32_t side, double drop_prob,
                       uint32_t n_long, uint64_t seed, int col_ptr_bytes,
                       uint64_t *nnz);
int bfsx_graph_info_get(bfsx_ctx *ctx, bfsx_graph_info *out);
/* Re-sorts every adjacency list of the resident graph in place (col_ptr is
 * unchanged): BFSX_ORDER_ID = ascending neighbour ids (the builder's and
 * scipy's CSC order), BFSX_ORDER_HUB_FIRST = neighbours by degree
 * descending, ties by id (sets BFSX_GRAPH_HUB_FIRST).  Either drops a
 * bottom-up copy.  BFSX_ORDER_PULL_HUB_FIRST keeps the lists as they are for
 * top-down steps and builds a second, hub-first copy of row_idx (4 bytes per
 * stored edge of HBM) that only bottom-up steps read (sets
 * BFSX_GRAPH_PULL_COPY; a new graph drops it).  Levels do not change;
 * bottom-up steps end their scans earlier on power-law graphs, while
 * top-down steps keep the id order's sequential sweep of the visited
 * bitmap.  A layout step of this engine: the reference keeps lists in file
 * order (src/graph_load.hpp:23-131). */
#define BFSX_ORDER_ID 0
#define BFSX_ORDER_HUB_FIRST 1
#define BFSX_ORDER_PULL_HUB_FIRST 2
int bfsx_order_adjacency(bfsx_ctx *ctx, int order);
/* Copies the resident graph to host buffers of n_cols+1 / nnz entries. */
int bfsx_download_csc(bfsx_ctx *ctx, void *col_ptr, uint32_t *row_idx);
/* Roots: mode 0 = seeded, among vertices of degree >= 1; mode 1 = seeded,
 * among the vertices of the largest connected component (symmetric graphs);
 * mode 2 = the n_roots vertices of largest degree, ties to the smaller id
 * (the seed is unused) -- scripts/rootNodes.py:133-159, which takes the top
 * 5 of networkx's degree (the number of neighbours on a simple symmetric
 * graph) with a stable sort.  Replaces the reference's offline root script. */
int bfsx_pick_roots(bfsx_ctx *ctx, uint32_t n_roots, uint64_t seed, int mode,
                    uint32_t *roots_out, uint32_t *n_found);

/* ---- search ------------------------------------------------------------- */
/* One BFS from root.  levels_out: n_cols int32 (-1 = unreached) or NULL;
 * parents_out: n_cols uint32 (0xFFFFFFFF = unreached, root -> root) or NULL;
 * stats: optional. */
int bfsx_run(bfsx_ctx *ctx, uint32_t root, int strategy, uint32_t run_flags,
             int32_t *levels_out, uint32_t *parents_out, bfsx_run_stats *stats);
/* n_roots searches back to back; levels_out is n_roots x n_cols (or NULL),
 * stats is n_roots entries (or NULL). */
int bfsx_run_batch(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots,
                   int strategy, uint32_t run_flags, int32_t *levels_out,
                   bfsx_run_stats *stats);
/* Per-level records of the last run made with BFSX_RUN_TRACE or
 * BFSX_RUN_STEP_TIMES (the latter fills level, choice, frontier_vertices,
 * discovered, us, pull_scanned and kernel_us). */
int bfsx_last_trace(bfsx_ctx *ctx, bfsx_level_record *out, uint32_t max_records,
                    uint32_t *n_records);
/* Level histogram of the last run, computed on the device: counts[l] =
 * vertices at level l, edges[l] = their out-degree sum (vertices < n_rows).
 * The reference prints and stores the same histogram (src/host.cpp:27-54,
 * :205-208, "LevelsGPU").  *n_levels = max level + 1. */
int bfsx_level_histogram(bfsx_ctx *ctx, uint64_t *counts, uint64_t *edges,
                         uint32_t max_levels, uint32_t *n_levels);
/* Device-to-device copy of `bytes` through a plain streaming kernel; *ms is
 * its duration.  Used to calibrate HBM counters and bandwidth. */
int bfsx_stream_copy(bfsx_ctx *ctx, uint64_t bytes, int repeats, double *ms);

/* ---- batched searches (no counterpart in the reference) -----------------
 * Up to 256 searches advance together, one bit each in a per-vertex bitset.
 * sources[i] starts search bits[i] at level 0 (several sources may share a
 * bit: the level is then the distance to the nearest one).  levels_out is
 * n_bits x n_cols uint8, one row per search (0xFF = unreached), or NULL.
 * Fails with BFSX_ERR_OVERFLOW if a level would exceed 254. */
typedef struct bfsx_batch_stats {
  double ms;              /* level loop of the whole batch */
  uint64_t edges_count;   /* sum over the searches of their edgesCount */
  uint32_t n_levels;
  uint32_t push_levels;   /* levels run top-down */
  uint32_t pull_levels;   /* levels run bottom-up (symmetric graphs) */
} bfsx_batch_stats;
int bfsx_run_multi(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits,
                   uint32_t n_sources, uint32_t n_bits, uint8_t *levels_out,
                   bfsx_batch_stats *stats);
/* Builds n_graphs R-MAT graphs (quadrants 0.57/0.19/0.19/0.05) packed into
 * one symmetric CSC: graph g holds vertices [offsets[g], offsets[g+1]).
 * offsets has n_graphs + 1 entries. */
/* Autotuner of the batched search: coordinate descent over ms_relabel,
 * ms_push_divisor, ms_pull_hub and ms_graph_levels, scoring the aggregate
 * throughput of bfsx_run_multi(sources, bits) on the resident graph; installs
 * and returns the best parameters like bfsx_autotune. */
int bfsx_autotune_multi(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits,
                        uint32_t n_sources, uint32_t n_bits, uint32_t passes, bfsx_params *best,
                        double *gteps_before, double *gteps_after);
int bfsx_generate_rmat_union(bfsx_ctx *ctx, uint32_t n_graphs, const int *scales,
                             const int *edge_factors, const uint64_t *seeds_graph,
                             const uint64_t *seeds_perm, int col_ptr_bytes,
                             uint32_t *offsets, uint64_t *nnz);

/* ---- learned per-level scheduler (no counterpart in the reference) ------
 * An FP32 MLP 8 -> 32 -> 32 -> 4 (ReLU) predicts log2(device microseconds)
 * of each step kind (LLB, GLB, SLB, PULL) from frontier statistics; with
 * strategy BFSX_LEARNED each level after the root level (one vertex, which
 * runs LLB) runs the kind with the smallest prediction (PULL only on
 * symmetric graphs); the decision is made on the device at the end of the
 * previous level, and a level whose frontier is within a factor 1.5 of the
 * one the last decision saw, after the same kind of step, keeps that
 * decision.
 * Without weights, or if the network returns NaN, BFSX_LEARNED runs fixed
 * LLB.  Parameters are 1476
 * floats: W1[32][8], b1[32], W2[32][32], b2[32], W3[4][32], b3[4]. */
#define BFSX_SCHED_PARAMS 1476
/* Loads scheduler weights (n == BFSX_SCHED_PARAMS); w == NULL unloads. */
int bfsx_sched_set_weights(bfsx_ctx *ctx, const float *w, uint32_t n);
/* Batch forward on the GPU: X is n x 8, Y is n x 4 (row-major). */
int bfsx_mlp_forward(bfsx_ctx *ctx, const float *X, uint32_t n, const float *W, float *Y);
/* Loss sum_ij M_ij (Y_ij - T_ij)^2 / sum M and its gradient (1476 floats)
 * computed by the GPU forward/backward kernels. */
int bfsx_mlp_loss_grad(bfsx_ctx *ctx, const float *X, const float *T, const float *M,
                       uint32_t n, const float *W, float *grad, float *loss);
/* Full-batch Adam on the GPU for `epochs` steps; W is updated in place;
 * loss_curve (optional) receives the loss before each step. */
int bfsx_mlp_train(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
                   float *W, uint32_t epochs, float lr, float *loss_curve);
/* Records training rows on the resident graph: for every level of a search
 * from each root, the 8 scheduler inputs (X, n x 8) and, for every step kind,
 * log2 of its device time from the same saved state (T, n x 4; M marks the
 * kinds that were allowed).  The search continues with the fastest kind, or
 * with a random one with probability epsilon.  Each kind is timed `repeats`
 * times (minimum kept). */
int bfsx_sched_record(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots,
                      uint32_t repeats, float epsilon, uint64_t seed, float *X, float *T,
                      float *M, uint32_t max_rows, uint32_t *n_rows);

/* ---- launch shape ------------------------------------------------------- */
int bfsx_get_params(bfsx_ctx *ctx, bfsx_params *out);
int bfsx_set_params(bfsx_ctx *ctx, const bfsx_params *in);
/* Autotuner (replaces the reference's hand-set SM_FACTOR and fixed
 * work-group size): coordinate descent over the launch-shape parameters that
 * `strategy` uses (BFSX_LEARNED tunes those of every step kind), scoring
 * each candidate by the harmonic-mean throughput of searches from `roots`
 * on the resident graph; `passes` sweeps over the parameters (1-4).  The
 * best parameters are installed in the context and copied to *best;
 * *gteps_before / *gteps_after (optional) report the score of the starting
 * and the final parameters. */
int bfsx_autotune(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots, int strategy,
                  uint32_t passes, bfsx_params *best, double *gteps_before,
                  double *gteps_after);

/* ---- allocation-free search (no counterpart in the reference) -----------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  One BFS
 * from `root` over a CSC graph the caller already holds in device memory
 * (col_ptr: n_cols+1 entries of col_ptr_bytes, row_idx: nnz entries; only
 * vertices < n_rows are expanded), writing int32 levels (-1 unreached) to the
 * caller's device array `levels` (n_cols entries); *n_levels (optional, host)
 * receives the number of levels, as bfsx_run's stats.n_levels.  Levels equal
 * bfsx_run's.  The call allocates nothing: the search state (visited bitmap,
 * two frontier queues, counters; BFSX_NOALLOC_SCRATCH_BYTES) lives in a fixed
 * device array of the library, one per GPU, held until the process ends, so
 * graphs are limited to BFSX_NOALLOC_MAX_VERTICES vertices (larger ones are
 * rejected with BFSX_ERR_ARG).  Every call re-initialises that state; calls
 * that overlap on one GPU race on it.  Kernels run on `stream` (a
 * hipStream_t; NULL = the null stream) and the call returns when the search
 * has finished; the calling thread's current device is restored.  A
 * col_ptr or row_idx entry outside the graph is skipped by the kernels and
 * reported as BFSX_ERR_ARG.  The message of the last failure on the calling
 * thread: bfsx_last_error(NULL). */
#define BFSX_NOALLOC_MAX_VERTICES (1u << 22)
#define BFSX_NOALLOC_SCRATCH_BYTES 34079488u
int bfsx_run_noalloc(int device, void *stream, const void *col_ptr, int col_ptr_bytes,
                     const uint32_t *row_idx, uint32_t n_cols, uint32_t n_rows, uint64_t nnz,
                     uint32_t root, int32_t *levels, uint32_t *n_levels);

#ifdef __cplusplus
}
#endif
#endif /* BFSX_H */
