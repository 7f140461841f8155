// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "bfsx.h"

namespace bfsx {

// Device-side control block of a search: frontier sizes of the two queue
// buffers, their degree sums, the GLB work-queue cursor and running totals.
// Lives in device memory; a pinned host mirror receives it once per level.
struct LevelCtl {
  uint32_t q_count[2];             // entries in queue buffer 0 / 1
  uint32_t level;                  // level of the frontier being expanded
  uint32_t overflow;               // set when a result overflows its type
  unsigned long long q_edges[2];   // sum of out-degree of queue 0 / 1
  unsigned long long glb_total;    // edges of the current GLB frontier
  uint32_t glb_cursor;             // GLB atomic work-queue head
  uint32_t glb_chunks;             // GLB chunk count
  unsigned long long visited_vertices;
  unsigned long long visited_edges;
  unsigned long long discovered;   // vertices found by the last step
  unsigned long long t_begin;      // device wall clock at step start
  unsigned long long t_end;        // device wall clock at step end
  uint32_t choice;                 // step kind picked for the next level
  uint32_t done;                   // frontier became empty
  float sched_out[4];              // scheduler outputs of the last decision
  float sched_in[8];               // scheduler inputs of the last decision
  uint32_t kind_count[4];          // levels expanded by each step kind
  uint32_t conv;                   // slot cursor of a bitmap-to-queue conversion
  unsigned long long examined;     // adjacency entries read by bottom-up steps (running)
  // frontier of the last network evaluation (a level whose frontier is
  // within a factor 1.5 of it, after the same kind of step, keeps that
  // decision)
  unsigned long long dec_EF;
  uint32_t dec_F, dec_pull, dec_valid;
  uint32_t decisions;  // network evaluations so far (sched_in/out hold the last)
  // batched searches: hub-list length of a top-down level, and whether the
  // previous level ran bottom-up (choice holds this level's direction)
  uint32_t ms_hub_count, ms_prev_pull;
  // grid-wide barrier of the device loop: arrivals, and the generation the
  // waiting work-groups watch (0 at search start; every barrier leaves the
  // count at 0)
  uint32_t bar_count;
  uint32_t pad_[30];  // bar_gen on another 128-byte line: polls do not hit the arrivals' line
  uint32_t bar_gen;
  uint32_t reserved;
  // Level ends so far on this context: only the level-end kernel changes it
  // (+1 each), so the tag it puts on the host-mapped words never repeats;
  // every other writer of the block keeps it (CTL_RESET_BYTES).
  uint32_t seq;
};
static_assert(offsetof(LevelCtl, seq) + sizeof(uint32_t) == sizeof(LevelCtl),
              "seq is the last field: resets and restores cover the bytes before it");
constexpr size_t CTL_RESET_BYTES = offsetof(LevelCtl, seq);

// What the host reads at a level end without synchronising the stream: four
// 8-byte words in host-mapped memory, written by the level-end kernel.
// Device writes to host memory can reach it out of order even across a
// system-scope fence, so every word carries the level end's sequence number
// and the host takes the words only once all four carry the number it waits
// for (an 8-byte aligned store arrives whole).
//   w[0] = seq << 32 | frontier size (count of the produced queue)
//   w[1] = (seq & 0xFFFF) << 48 | frontier edges (48 bits)
//   w[2] = (seq & 0xFFFF) << 48 | edges of the visited vertices (48 bits)
//   w[3] = seq << 32 | next step kind (8 bits) | done << 8 | level << 9
//          (the level of the next frontier, 23 bits)
// Two such sets alternate by the parity of seq: a level end launched ahead
// of the host's read of the previous one (a speculative step) writes the
// other set, so it cannot overwrite words the host has not read yet.
struct LevelMirror {
  unsigned long long w[4];
};
constexpr int MIRROR_SLOTS = 2;
constexpr unsigned long long MIRROR_48 = (1ull << 48) - 1;
constexpr uint32_t MIRROR_LEVEL_MASK = (1u << 23) - 1;

// Whether a top-down level marks and compacts or discovers exact-once.  A
// level marks from hi_E frontier edges on, and from lo_E on when its
// frontier averages at least MARK_MIN_AVG_DEGREE edges per vertex.  Many
// edges per frontier vertex (hub lists) send many edges to the same new
// vertices, whose exact-once atomics then serialise; few (a grid's levels)
// find mostly distinct new vertices, and their appends cost less than the
// compaction's pass over all n marks.
constexpr uint32_t MARK_MIN_AVG_DEGREE = 8;

struct MarkRule {
  unsigned long long lo_E, hi_E;
  __host__ __device__ bool marks(unsigned long long EF, uint32_t F) const {
    return EF >= hi_E || (EF >= lo_E && EF >= (unsigned long long)MARK_MIN_AVG_DEGREE * F);
  }
};

// The host's copy of those values (from the mirror, or from a full copy of
// the control block).
struct LevelView {
  uint32_t F = 0;
  unsigned long long EF = 0, VE = 0;
  uint32_t choice = 0, done = 0;
  uint32_t level = 0;  // level of the next frontier (23 bits)
};

// Bitmaps hold whole 1024-vertex chunks (32 words) so that a wave can own a
// chunk; mark arrays hold one byte per vertex, padded the same way.
__host__ __device__ inline uint64_t bitmap_words(uint64_t n) { return ((n + 1023) / 1024) * 32; }

// A wave of the bottom-up step takes 2048 vertices (64 visited-bitmap words)
// at a time.
constexpr uint32_t PULL_CHUNK = 2048;
// Largest frontier summary of the bottom-up step (words; 32 KiB).
constexpr uint32_t PULL_SUM_WORDS_MAX = 8192;
constexpr uint32_t DEG16_ESCAPE = 0xFFFFu;
// fd[v].x with this bit set names a hub by its index (graphs below 2^31
// vertices); at most PULL_HUBS_MAX hubs (a 32 KiB bitmap).
constexpr uint32_t PULL_HUB_FLAG = 0x80000000u;
constexpr uint32_t PULL_HUBS_MAX = 1u << 18;
__host__ __device__ inline uint64_t mark_bytes(uint64_t n) { return ((n + 1023) / 1024) * 1024; }

// Growable device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace bfsx

struct bfsx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  bool coop_ok = false;            // cooperative launches supported
  uint32_t lds_per_block = 65536;  // LDS bytes one work-group may allocate
  uint32_t lds_per_cu = 65536;     // LDS bytes of one CU
  int wall_clock_khz = 100000;
  std::string name;
  uint64_t hbm_bytes = 0;
  std::string err;

  // resident graph
  bool has_graph = false;
  uint32_t n_rows = 0, n_cols = 0;
  uint64_t nnz = 0;
  int ptr_bytes = 4;
  uint32_t gflags = 0;
  uint32_t max_degree = 0, n_isolated = 0;
  bfsx::DevBuf col_ptr, row_idx, iso_bm;
  // min(out-degree, DEG16_ESCAPE) per vertex, padded with zeros to whole
  // visited-bitmap words: the degree sums of new frontiers read 2 bytes per
  // vertex instead of two col_ptr entries (col_ptr only for the escapes)
  bfsx::DevBuf deg16;
  bfsx::DevBuf row_idx_pull;  // hub-first copy read by bottom-up steps (optional)
  // With it: the first (highest-degree) neighbour (uint32, pull_fd) and the
  // degree (uint16, min(degree, DEG16_ESCAPE), pull_fdd) of every vertex with
  // neighbours, in vertex order and nothing for the isolated ones (a
  // bottom-up step reaches only vertices with neighbours: the isolated ones
  // start visited).  The entry of v sits at pull_crank[v / PULL_CHUNK] plus
  // the vertices with neighbours before v in its PULL_CHUNK-vertex chunk
  // (from the isolated-vertex bitmap); one padding entry follows.  While the
  // arrays are built, pull_fd holds uint2 {first neighbour, degree} per vertex.
  bfsx::DevBuf pull_fd, pull_fdd, pull_crank;
  // The pull_hubs vertices of largest degree ("hubs") are named in pull_fd by
  // their index: fd[v].x = PULL_HUB_FLAG | h for the hub pull_hub[h].  Before
  // every bottom-up step their frontier bits are gathered into pull_hbm
  // (pull_hubs bits), a bitmap small enough to stay in L2.
  bfsx::DevBuf pull_hub, pull_hbm;
  uint32_t pull_hubs = 0;  // a multiple of 64; 0: pull_fd holds plain ids
  // Top-down copy of row_idx in which every neighbour among the hot_n
  // vertices of largest degree is named by its index (PULL_HUB_FLAG | h for
  // hot_id[h]); heavy LLB/SLB levels read it and probe those neighbours in
  // hot_vis (hot_n bits, rebuilt from the visited bitmap at each such level,
  // small enough to stay in L2).  Built with the bottom-up copy.
  bfsx::DevBuf row_idx_hot, hot_id, hot_vis;
  uint32_t hot_n = 0;

  // per-search state, sized for the resident graph
  bfsx::DevBuf levels, parents, vis, fbm[2], queue[2], ctl, mark;
  bfsx::DevBuf glb_scan, glb_start, glb_tile, glb_chunk_lo;
  bfsx::DevBuf fsum;     // bottom-up step: frontier summary bitmap (PULL_SUM_WORDS_MAX)
  bfsx::DevBuf bkt, bkt_cnt, bkt_tot, bkt_base;  // GLB bucketed discovery
  bfsx::DevBuf scratch;  // generic scratch (reductions, root picking)
  bfsx::LevelCtl *h_ctl = nullptr;        // pinned: full copies of the control block
  bfsx::LevelMirror *h_mirror = nullptr;  // pinned, host-mapped: tagged level-end words (2 sets)
  bfsx::LevelMirror *d_mirror = nullptr;  // device address of h_mirror (k_level_end writes it)
  bfsx::LevelView view;                   // the values of the last level end the host read
  bool mark_dirty = false;          // a mark step did not reach its compaction

  // learned scheduler weights and the snapshot used to time alternatives
  bfsx::DevBuf sched_w;
  bool sched_loaded = false;
  bfsx::DevBuf snap_levels, snap_vis, snap_ctl;

  bfsx_params params{};
  std::vector<bfsx_level_record> trace;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_step0 = nullptr, ev_step1 = nullptr;
  hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr;  // main kernel of a timed step

  // batched multi-source search state
  bfsx::DevBuf ms_seen, ms_cur, ms_nxt, ms_lv, ms_hubs, ms_phubs, ms_act;
  // the graph relabelled by degree for batched searches (made on first use)
  bfsx::DevBuf ms_rcp, ms_rri, ms_rank, ms_order;
  bool ms_relabelled = false;
  // captured level loop of the batched search, rebuilt when its key changes
  hipGraph_t ms_graph = nullptr;
  hipGraphExec_t ms_exec = nullptr;
  uint64_t ms_key = 0;

  // captured level loop (BFSX_RUN_GRAPH_LOOP), rebuilt when its key changes
  hipGraph_t loop_graph = nullptr;
  hipGraphExec_t loop_exec = nullptr;
  uint64_t loop_key = 0;
};

namespace bfsx {

// Error plumbing: every entry point catches and converts to a status code.
struct Error {
  int status;
  std::string msg;
};
[[noreturn]] void fail(int status, const std::string &msg);
void check_hip(hipError_t e, const char *what);
#define BFSX_HIP(call) ::bfsx::check_hip((call), #call)

void ensure(DevBuf &b, size_t bytes, hipStream_t s);
void release(DevBuf &b);

// ---- graph construction (bfsx_gen.hip) -----------------------------------
// Builds a symmetric CSC in ctx from n undirected pairs (u[i], v[i]) resident
// on the device; u/v buffers are consumed.
void build_symmetric_csc(bfsx_ctx *ctx, uint32_t n_vertices, uint32_t *d_u,
                         uint32_t *d_v, uint64_t m, int ptr_bytes);
void gen_rmat_pairs(bfsx_ctx *ctx, int scale, uint64_t m, uint32_t tA,
                    uint32_t tAB, uint32_t tABC, uint64_t seed_graph,
                    uint64_t seed_perm, uint32_t *d_u, uint32_t *d_v);
uint64_t gen_grid_pairs(bfsx_ctx *ctx, uint32_t side, uint32_t t_drop,
                        uint32_t n_long, uint64_t seed, uint32_t *d_u,
                        uint32_t *d_v);
void graph_summary(bfsx_ctx *ctx);  // max degree, isolated bitmap/count
// re-sorts every adjacency list by neighbour id or hub first (needs graph_summary)
// Sorts every adjacency list (id order, or hub first) into `out`, which is
// ctx->row_idx (in place) or a copy of the same size.
void order_adjacency(bfsx_ctx *ctx, bool hub_first, uint32_t *out);
// fd[v] = {row_idx[col_ptr[v]] (~0u without neighbours), degree} for v < n_rows
// (degrees below 2^32).
void first_neighbours(bfsx_ctx *ctx, const uint32_t *row_idx, uint2 *fd);
// Names the first neighbours that are among the graph's top-degree vertices
// by their index in ctx->pull_hub (sets pull_hub, pull_hbm, pull_hubs); needs
// ctx->pull_fd (one entry per vertex) and n_cols < 2^31.
void index_hub_neighbours(bfsx_ctx *ctx);
// Drops the isolated vertices' entries from ctx->pull_fd (one uint2 entry per
// vertex before), splits it into ctx->pull_fd (first neighbours) and
// ctx->pull_fdd (16-bit degrees), and builds ctx->pull_crank (needs the
// isolated bitmap).
void compact_first_neighbours(bfsx_ctx *ctx);
// Builds row_idx_hot / hot_id / hot_vis / hot_n for graphs of
// HOT_MIN_VERTICES to 2^31 vertices (otherwise leaves hot_n = 0).
constexpr uint32_t HOT_MIN_VERTICES = 1u << 24;
void build_hot_copy(bfsx_ctx *ctx);
// The graph with vertices renamed by (degree descending, id ascending):
// rank[v] = new id of v, order[i] = old id of new vertex i, new_cp / new_ri
// the relabelled CSC (lists ascending in new ids), col_ptr of the graph's
// width.  Needs n_rows == n_cols.
void relabel_by_degree(bfsx_ctx *ctx, uint32_t *rank, uint32_t *order, void *new_cp,
                       uint32_t *new_ri);
// d_out[0..min(k, n_rows)) = vertices of largest out-degree, ties by id
void top_degree_vertices(bfsx_ctx *ctx, uint32_t k, uint32_t *d_out);
void root_candidates(bfsx_ctx *ctx, uint64_t seed, uint64_t first,
                     uint32_t count, const int32_t *d_component_levels,
                     uint32_t *d_out_vertex, uint8_t *d_out_ok);

// ---- search kernels (bfsx_bfs.hip) ---------------------------------------
struct StepArgs {
  const void *col_ptr;
  const uint32_t *row_idx;
  uint32_t n_rows, n_cols;
  const uint32_t *q_in;
  uint32_t q_in_idx;       // which LevelCtl counter describes q_in
  uint32_t F;              // frontier size (host-known)
  uint32_t *q_out;
  uint32_t q_out_idx;
  uint32_t *vis;
  int32_t *levels;
  uint32_t *parents;       // nullable
  uint32_t *fbm_cur;       // frontier bitmap of this level (pull)
  uint32_t *fbm_next;      // next frontier bitmap (pull writes it)
  LevelCtl *ctl;
  int32_t next_level;
  bool track_degrees;      // accumulate degree sum of discovered vertices
  uint8_t *mark;           // non-null: mark-then-compact step
  bool count_examined;     // bottom-up: add adjacency entries read to ctl->examined
  hipEvent_t ev_k0, ev_k1; // non-null: recorded around the step's main kernel
  unsigned long long EF;   // frontier edges (out-degree sum of q_in)
  bool bucket_ok;          // a marking GLB step may use bucketed discovery
  bool hot_ok;             // a marking LLB/SLB step may read the hot copy (most
                           // hot vertices visited: their probes then stop there)
};

void init_search(bfsx_ctx *ctx, uint32_t root, bool symmetric_iso, uint32_t *parents);
void step_llb(bfsx_ctx *ctx, const StepArgs &a, bool strided);
// Returns true when the step discovered with buckets (heavy level, no
// parents): the next queue, levels, visited bits and (if a.fbm_next) the
// next frontier bitmap are then final and no compaction follows.
bool step_glb(bfsx_ctx *ctx, const StepArgs &a);
// speculative: launched before the previous level's end has decided this
// level's kind (see SpecGuard); it runs only if that decision is PULL
void step_pull(bfsx_ctx *ctx, const StepArgs &a, bool speculative = false);
void bitmap_to_queue(bfsx_ctx *ctx, const uint32_t *bm, uint32_t *q);
void queue_to_fbm(bfsx_ctx *ctx, const uint32_t *q, uint32_t F, uint32_t *fbm);
void compact(bfsx_ctx *ctx, const StepArgs &s, uint8_t *mark, uint32_t *fbm_next);
// Level bookkeeping; with decide_allow != 0 the learned scheduler also picks
// the next level's step kind (among the allowed ones) in the same launch,
// and with reuse it keeps the last decision while the frontier stays close
// to the one that decision was computed for.
// spec_kinds != 0: launched ahead of the previous level's decision; runs
// only if that decision is one of these kinds (bit k for kind k) and the
// frontier is not empty and (spec_heavy) not one that rule marks.
void level_end(bfsx_ctx *ctx, uint32_t consumed_idx, uint32_t produced_idx,
               bool stamp, int kind, uint32_t decide_allow = 0, bool reuse = false,
               uint32_t spec_kinds = 0, const MarkRule *spec_heavy = nullptr);
// mark: which top-down levels mark and compact (needs track: the edge
// counts are kept in the control block).  speculative: launched ahead of
// the previous level's decision; the step then skips frontiers `mark` marks
// (no compaction follows) besides kinds other than LLB and SLB.
void dev_level(bfsx_ctx *ctx, uint32_t cur, int fixed, bool track, MarkRule mark,
               bool speculative = false);
void level_begin(bfsx_ctx *ctx);
// Runs consecutive thin top-down levels (levels `mark` does not mark,
// kind LLB or SLB: `fixed`, or the scheduler's choice when fixed < 0) in one
// cooperative launch, from the frontier in queue `cur`; the launch ends with
// one set of level-end words (ctx->d_mirror) carrying the level reached.
// False when the device cannot launch cooperatively.
bool coop_levels(bfsx_ctx *ctx, uint32_t cur, int fixed, uint32_t allow, bool learned,
                 MarkRule mark, uint32_t max_levels);
// Waits for the level end with sequence number `seq` (the tagged host-mapped
// words) and leaves its values in ctx->view; `produced` = the queue buffer
// the level filled (for the synchronising fallback).
// After a device loop (level0 = the level at its launch) the buffer is
// `produced` flipped once per level it ran.
void read_level_end(bfsx_ctx *ctx, uint32_t seq, uint32_t produced, uint32_t level0 = ~0u);
void reached_summary(bfsx_ctx *ctx, unsigned long long *d_out);
void level_histogram(bfsx_ctx *ctx, uint32_t max_levels, unsigned long long *d_counts,
                     unsigned long long *d_edges);
void stream_copy(bfsx_ctx *ctx, const void *src, void *dst, uint64_t bytes);

// ---- batched multi-source search (bfsx_msbfs.hip) -------------------------
struct MsResult {
  double ms = 0;
  uint32_t n_levels = 0, push_levels = 0, pull_levels = 0;
  unsigned long long edges = 0;
};
MsResult ms_run(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits, uint32_t n_src,
                uint32_t n_bits, uint8_t *levels_out);
void ms_drop(bfsx_ctx *ctx);  // the relabelled graph and the captured loop
void gen_rmat_pairs_offset(bfsx_ctx *ctx, int scale, uint64_t m, uint32_t tA, uint32_t tAB,
                           uint32_t tABC, uint64_t seed_graph, uint64_t seed_perm,
                           uint32_t offset, uint32_t *d_u, uint32_t *d_v);

// ---- allocation-free search (bfsx_noalloc.hip) ----------------------------
const char *noalloc_last_error();  // the calling thread's last failure message

// ---- learned scheduler (bfsx_sched.hip) -----------------------------------
// Features of the frontier in queue `idx` -> MLP -> ctl->choice.
void decide(bfsx_ctx *ctx, uint32_t idx, bool first, uint32_t allow_mask);
void mlp_forward(bfsx_ctx *ctx, const float *X, uint32_t n, const float *W, float *Y);
void mlp_loss_grad(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
                   const float *W, float *grad, float *loss);
void mlp_train(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
               float *W, uint32_t epochs, float lr, float *loss_curve);

}  // namespace bfsx
