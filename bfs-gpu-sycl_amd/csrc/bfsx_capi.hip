// Engine entry points: context, resident graph, the level loop, C ABI.
//
// The level loop replaces GPURun's (src/run_bfs_gpu.hpp:679-728): per level
// the host launches the chosen step, one small bookkeeping kernel, and reads
// the level control block back (the reference reads the 4-byte frontier
// count, :713-715).  Device memory for the graph and all search state is
// allocated once per graph and reused across runs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <new>

#include "bfsx_internal.hpp"

namespace bfsx {

void fail(int status, const std::string &msg) { throw Error{status, msg}; }

void check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  const int st = (e == hipErrorOutOfMemory) ? BFSX_ERR_NOMEM : BFSX_ERR_DEVICE;
  fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

void ensure(DevBuf &b, size_t bytes, hipStream_t s) {
  (void)s;
  if (bytes == 0) bytes = 1;
  if (b.p && b.bytes >= bytes) return;
  if (b.p) {
    BFSX_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  BFSX_HIP(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
}

void release(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

static void set_device(bfsx_ctx *ctx) { BFSX_HIP(hipSetDevice(ctx->device)); }

void drop_loop_graph(bfsx_ctx *ctx);

static void free_search_state(bfsx_ctx *ctx) {
  drop_loop_graph(ctx);  // the captured loop holds pointers into this state
  ms_drop(ctx);
  for (DevBuf *b : {&ctx->levels, &ctx->parents, &ctx->vis, &ctx->fbm[0], &ctx->fbm[1],
                    &ctx->queue[0], &ctx->queue[1], &ctx->glb_scan, &ctx->glb_start,
                    &ctx->glb_tile, &ctx->glb_chunk_lo, &ctx->mark, &ctx->fsum, &ctx->bkt,
                    &ctx->bkt_cnt, &ctx->bkt_tot, &ctx->bkt_base, &ctx->ms_seen,
                    &ctx->ms_cur, &ctx->ms_nxt, &ctx->ms_lv, &ctx->ms_hubs, &ctx->ms_phubs,
                    &ctx->ms_act})
    release(*b);
}

static void alloc_search_state(bfsx_ctx *ctx) {
  hipStream_t s = ctx->stream;
  const uint64_t n = ctx->n_cols;
  ensure(ctx->levels, sizeof(int32_t) * n, s);
  ensure(ctx->vis, sizeof(uint32_t) * bitmap_words(n), s);
  ensure(ctx->queue[0], sizeof(uint32_t) * (n + 1), s);
  ensure(ctx->queue[1], sizeof(uint32_t) * (n + 1), s);
  ensure(ctx->mark, mark_bytes(n), s);
  BFSX_HIP(hipMemsetAsync(ctx->mark.p, 0, mark_bytes(n), s));
  BFSX_HIP(hipStreamSynchronize(s));
}

static void ensure_pull_state(bfsx_ctx *ctx) {
  const uint64_t w = bitmap_words(ctx->n_cols);
  ensure(ctx->fbm[0], sizeof(uint32_t) * w, ctx->stream);
  ensure(ctx->fbm[1], sizeof(uint32_t) * w, ctx->stream);
  ensure(ctx->fsum, sizeof(uint32_t) * bfsx::PULL_SUM_WORDS_MAX, ctx->stream);
}

// the {first neighbour, degree} array and its hub index, and the top-down
// copy with hot neighbours named by index
static void drop_pull_fd(bfsx_ctx *ctx) {
  release(ctx->pull_fd);
  release(ctx->pull_fdd);
  release(ctx->pull_crank);
  release(ctx->pull_hub);
  release(ctx->pull_hbm);
  ctx->pull_hubs = 0;
  release(ctx->row_idx_hot);
  release(ctx->hot_id);
  release(ctx->hot_vis);
  ctx->hot_n = 0;
}

static void after_graph_change(bfsx_ctx *ctx) {
  // edge counts travel in 48-bit fields of the level-end words
  if (ctx->nnz > MIRROR_48) fail(BFSX_ERR_ARG, "graphs above 2^48 stored edges are not supported");
  release(ctx->row_idx_pull);  // a copy of the old lists
  drop_pull_fd(ctx);
  ctx->gflags &= ~BFSX_GRAPH_PULL_COPY;
  free_search_state(ctx);
  graph_summary(ctx);
  alloc_search_state(ctx);
}

static bfsx_params default_params() {
  bfsx_params p{};
  p.llb_wg = 256;
  p.glb_wg = 256;
  p.glb_blocks_per_cu = 8;
  p.glb_chunk = 4096;
  p.hub_degree = 0;
  p.pull_wg = 512;
  p.graph_levels = 8;
  p.mark_ratio_q10 = 4;
  p.bucket_ratio_q10 = 512;
  p.ms_relabel = 1;
  p.ms_push_divisor = 14;
  p.ms_pull_hub = 256;
  p.ms_graph_levels = 8;
  p.coop_groups = 256;
  p.speculate = 1;
  return p;
}

static void validate_params(const bfsx_params &p) {
  auto pow2_in = [](uint32_t v, uint32_t lo, uint32_t hi) {
    return v >= lo && v <= hi && (v & (v - 1)) == 0;
  };
  if (!pow2_in(p.llb_wg, 64, 1024)) fail(BFSX_ERR_ARG, "llb_wg must be a power of two in [64,1024]");
  if (!pow2_in(p.glb_wg, 128, 1024)) fail(BFSX_ERR_ARG, "glb_wg must be a power of two in [128,1024]");
  if (!pow2_in(p.pull_wg, 64, 512)) fail(BFSX_ERR_ARG, "pull_wg must be a power of two in [64,512]");
  if (p.glb_blocks_per_cu < 1 || p.glb_blocks_per_cu > 64) fail(BFSX_ERR_ARG, "glb_blocks_per_cu must be in [1,64]");
  if (p.glb_chunk < 256 || p.glb_chunk > (1u << 22)) fail(BFSX_ERR_ARG, "glb_chunk must be in [256, 2^22]");
  if (p.graph_levels < 1 || p.graph_levels > 64) fail(BFSX_ERR_ARG, "graph_levels must be in [1,64]");
  if (p.ms_relabel > 1) fail(BFSX_ERR_ARG, "ms_relabel must be 0 or 1");
  if (p.coop_groups < 1 || p.coop_groups > 4096)
    fail(BFSX_ERR_ARG, "coop_groups must be in [1,4096]");
  if (p.speculate > 1) fail(BFSX_ERR_ARG, "speculate must be 0 or 1");
  if (p.ms_push_divisor < 1 || p.ms_push_divisor > 1024) fail(BFSX_ERR_ARG, "ms_push_divisor must be in [1,1024]");
  if (p.ms_pull_hub < 32 || p.ms_pull_hub > 65536) fail(BFSX_ERR_ARG, "ms_pull_hub must be in [32,65536]");
  if (p.ms_graph_levels < 2 || p.ms_graph_levels > 64 || (p.ms_graph_levels & 1))
    fail(BFSX_ERR_ARG, "ms_graph_levels must be even and in [2,64]");
}

// ---------------------------------------------------------------------------
// one search
// ---------------------------------------------------------------------------
struct RunOut {
  double ms = 0;
  uint32_t n_levels = 0;
  uint32_t steps[BFSX_N_STEP_KINDS] = {0, 0, 0, 0};
};

// Full copy of the control block into h_ctl (stream synchronised).
static void read_ctl(bfsx_ctx *ctx) {
  BFSX_HIP(hipMemcpyAsync(ctx->h_ctl, ctx->ctl.p, sizeof(LevelCtl), hipMemcpyDeviceToHost,
                          ctx->stream));
  BFSX_HIP(hipStreamSynchronize(ctx->stream));
}

// ctx->view from the full copy in h_ctl; `produced` = the queue buffer the
// last step filled.
static void view_from_ctl(bfsx_ctx *ctx, uint32_t produced) {
  const LevelCtl &c = *ctx->h_ctl;
  ctx->view.F = c.q_count[produced];
  ctx->view.EF = c.q_edges[produced];
  ctx->view.VE = c.visited_edges;
  ctx->view.choice = c.choice;
  ctx->view.done = c.done;
  ctx->view.level = c.level & MIRROR_LEVEL_MASK;
}

// After a launch that ends with k_level_end: that kernel writes the values
// the host needs into the host-mapped mirror, every word tagged with its
// sequence number, so the host polls the words (no copy, no stream
// synchronisation) until all four carry `seq` -- a word of the previous
// level end next to a new one means the new ones have not all arrived.
// After a second without them, the stream is synchronised, which reports any
// device error, and the values come from a full copy of the control block.
void read_level_end(bfsx_ctx *ctx, uint32_t seq, uint32_t produced, uint32_t level0) {
  const volatile unsigned long long *w = ctx->h_mirror[seq & 1u].w;
  const unsigned long long t32 = seq, t16 = seq & 0xFFFFu;
  const auto t0 = std::chrono::steady_clock::now();
  auto take = [&]() {
    const unsigned long long w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    if (!((w0 >> 32) == t32 && (w1 >> 48) == t16 && (w2 >> 48) == t16 && (w3 >> 32) == t32))
      return false;
    ctx->view.F = (uint32_t)w0;
    ctx->view.EF = w1 & MIRROR_48;
    ctx->view.VE = w2 & MIRROR_48;
    ctx->view.choice = (uint32_t)(w3 & 0xFFu);
    ctx->view.done = (uint32_t)((w3 >> 8) & 1u);
    ctx->view.level = (uint32_t)((w3 >> 9) & MIRROR_LEVEL_MASK);
    return true;
  };
  for (uint32_t spin = 0;; spin++) {
    if (take()) return;
    if ((spin & 1023u) == 1023u &&
        std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
      read_ctl(ctx);  // synchronises the stream (reports device errors)
      // every write has landed now; with a speculative level end after
      // this one, the control block is already past it, but this level's
      // words are in their own set
      if (take()) return;
      if (ctx->h_ctl->seq != seq) fail(BFSX_ERR_DEVICE, "level end did not arrive");
      // after a device loop: the frontier's buffer follows the levels run
      if (level0 != ~0u) produced ^= (ctx->h_ctl->level - level0) & 1u;
      view_from_ctl(ctx, produced);
      return;
    }
  }
}

// Which top-down levels mark and compact (the others discover exact-once):
// from n_cols * mark_ratio_q10 / 1024 frontier edges when the frontier
// averages at least MARK_MIN_AVG_DEGREE edges per vertex, from n_cols / 16
// (or the ratio's count, if larger) whatever the frontier; ratio 0: every
// level, 0xFFFFFFFF: none.
static MarkRule mark_rule(const bfsx_ctx *ctx) {
  const uint32_t r = ctx->params.mark_ratio_q10;
  if (r == 0) return MarkRule{0, 0};
  if (r == 0xFFFFFFFFu) return MarkRule{~0ull, ~0ull};
  const unsigned long long lo = ((unsigned long long)ctx->n_cols * r + 1023) / 1024;
  const unsigned long long all = ((unsigned long long)ctx->n_cols + 15) / 16;
  return MarkRule{lo, lo > all ? lo : all};
}

static bool graph_symmetric(const bfsx_ctx *ctx) {
  return (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols;
}

// Host-side state of one level-synchronous search; the device side lives in
// ctx (levels, vis, queues, ctl).  One call of step() runs one level.
struct Walker {
  bfsx_ctx *ctx;
  uint32_t F = 1, cur = 0, level = 0;
  unsigned long long EF = 0;  // frontier edges (kept by the kernels: track is on)
  unsigned long long VE = 0;  // edges of the visited vertices
  int fbm_cur = 0;         // buffer holding the current frontier's bitmap
  bool fbm_valid = false;  // whether that bitmap is up to date
  bool queue_valid = true; // false after a bottom-up step (frontier only as bitmap)
  bool pull_possible = false;
  bool track = false;      // accumulate frontier degree sums
  bool stamp = false;      // device clock around each step
  bool count = false;      // count adjacency entries read by bottom-up steps
  uint32_t *par = nullptr;
  int last = BFSX_LLB;     // step kind of the last launch
  uint32_t decide_allow = 0;  // learned: kinds the next decision may pick
  bool decide_reuse = false;  // keep the last decision on a similar frontier

  // Launches the kernels of one level with step kind `choice`, then the
  // level bookkeeping kernel.  The caller reads the control block back and
  // calls advance().
  bool launch(int choice) {
    StepArgs a{};
    a.col_ptr = ctx->col_ptr.p;
    a.row_idx = ctx->row_idx.as<uint32_t>();
    a.n_rows = ctx->n_rows;
    a.n_cols = ctx->n_cols;
    a.q_in = ctx->queue[cur].as<uint32_t>();
    a.q_in_idx = cur;
    a.F = F;
    a.q_out = ctx->queue[cur ^ 1].as<uint32_t>();
    a.q_out_idx = cur ^ 1;
    a.vis = ctx->vis.as<uint32_t>();
    a.levels = ctx->levels.as<int32_t>();
    a.parents = par;
    a.ctl = ctx->ctl.as<LevelCtl>();
    a.next_level = (int32_t)level + 1;
    a.track_degrees = track;
    a.count_examined = count;
    a.ev_k0 = count ? ctx->ev_k0 : nullptr;
    a.ev_k1 = count ? ctx->ev_k1 : nullptr;
    a.EF = EF;
    const uint32_t br = ctx->params.bucket_ratio_q10;
    const unsigned long long UE = ctx->nnz > VE ? ctx->nnz - VE : 0ull;
    const MarkRule mr = mark_rule(ctx);
    const bool use_mark = choice != BFSX_PULL && mr.marks(EF, F);
    // bucketed GLB from n_cols/256 frontier edges on (below that its fixed
    // cost, about a pass over the visited bitmap, exceeds exact-once's
    // atomics) or wherever a marking step would run
    a.bucket_ok = br != 0xFFFFFFFFu && (double)UE * 1024.0 >= (double)ctx->nnz * br &&
                  (use_mark || EF * 256 >= ctx->n_cols) && mr.lo_E != ~0ull;
    a.mark = use_mark ? ctx->mark.as<uint8_t>() : nullptr;
    // the hot copy pays once the visited vertices hold half the edges
    a.hot_ok = VE * 2 >= ctx->nnz;
    // next frontier bitmap of a top-down step (kept for bottom-up steps)
    uint32_t *td_fbm_next = pull_possible ? ctx->fbm[fbm_cur ^ 1].as<uint32_t>() : nullptr;
    bool bucketed = false;
    if (stamp) level_begin(ctx);
    if (choice != BFSX_PULL && !queue_valid)
      bitmap_to_queue(ctx, ctx->fbm[fbm_cur].as<uint32_t>(), ctx->queue[cur].as<uint32_t>());
    bool new_fbm = false;
    switch (choice) {
      case BFSX_LLB: step_llb(ctx, a, false); break;
      case BFSX_SLB: step_llb(ctx, a, true); break;
      case BFSX_GLB:
        a.fbm_next = td_fbm_next;
        bucketed = step_glb(ctx, a);
        break;
      case BFSX_PULL:
        if (!pull_possible) fail(BFSX_ERR_UNSUPPORTED, "bottom-up step on a non-symmetric graph");
        if (ctx->row_idx_pull.p) a.row_idx = ctx->row_idx_pull.as<uint32_t>();
        // (the queue is valid whenever the bitmap is not: only a bottom-up
        // step leaves the queue stale, and it writes the bitmap)
        if (!fbm_valid) queue_to_fbm(ctx, a.q_in, F, ctx->fbm[fbm_cur].as<uint32_t>());
        a.fbm_cur = ctx->fbm[fbm_cur].as<uint32_t>();
        a.fbm_next = ctx->fbm[fbm_cur ^ 1].as<uint32_t>();
        step_pull(ctx, a);
        new_fbm = true;
        break;
      default: fail(BFSX_ERR_ARG, "unknown step kind");
    }
    if (bucketed) {
      new_fbm = td_fbm_next != nullptr;
    } else if (use_mark) {
      ctx->mark_dirty = true;
      compact(ctx, a, a.mark, td_fbm_next);
      new_fbm = td_fbm_next != nullptr;
    }
    level_end(ctx, cur, cur ^ 1, stamp, choice, decide_allow, decide_reuse);
    last = choice;
    return new_fbm;
  }

  // Launches, right after this level's kernels and before its end has been
  // read, the bottom-up step of the next level, guarded on the device: it
  // runs only if this level's end picks PULL for the next level and the
  // frontier is not empty, so that a run of bottom-up levels does not wait
  // for a host round trip per level.  Only after a step that wrote the next
  // frontier as a bitmap (bottom-up, or bucketed or marking top-down),
  // learned, untimed.
  void launch_speculative_pull() {
    StepArgs a{};
    a.col_ptr = ctx->col_ptr.p;
    a.row_idx = ctx->row_idx_pull.p ? ctx->row_idx_pull.as<uint32_t>() : ctx->row_idx.as<uint32_t>();
    a.n_rows = ctx->n_rows;
    a.n_cols = ctx->n_cols;
    a.q_in = ctx->queue[cur ^ 1].as<uint32_t>();
    a.q_in_idx = cur ^ 1;
    a.F = 0;  // not known yet
    a.q_out = ctx->queue[cur].as<uint32_t>();
    a.q_out_idx = cur;
    a.vis = ctx->vis.as<uint32_t>();
    a.levels = ctx->levels.as<int32_t>();
    a.parents = par;
    a.ctl = ctx->ctl.as<LevelCtl>();
    a.next_level = (int32_t)level + 2;
    a.track_degrees = track;
    a.count_examined = count;
    a.fbm_cur = ctx->fbm[fbm_cur ^ 1].as<uint32_t>();
    a.fbm_next = ctx->fbm[fbm_cur].as<uint32_t>();
    step_pull(ctx, a, true);
    level_end(ctx, cur ^ 1, cur, false, BFSX_PULL, decide_allow, decide_reuse, 1u << BFSX_PULL);
  }

  // The same after a top-down level, whose successor is often a thin
  // top-down level too (the tail of a search): the next level's LLB or SLB
  // step, sized on the device (the captured loop's kernel: frontier, level
  // and kind come from the control block), runs only if the decision is LLB
  // or SLB and the frontier is not empty and not one `mr` marks (a heavy
  // level runs the host-sized step).  That kernel records no parents: not
  // with a parent array.
  void launch_speculative_td(const MarkRule &mr) {
    dev_level(ctx, cur ^ 1, -1, true, mr, true);
    level_end(ctx, cur ^ 1, cur, false, -1, decide_allow, decide_reuse,
              (1u << BFSX_LLB) | (1u << BFSX_SLB), &mr);
  }

  // Host bookkeeping after the control block of the level has been read.
  void advance(bool new_fbm) {
    ctx->mark_dirty = false;
    if (new_fbm) fbm_cur ^= 1;
    fbm_valid = new_fbm;
    queue_valid = last != BFSX_PULL;
    F = ctx->view.F;
    EF = ctx->view.EF;
    VE = ctx->view.VE;
    cur ^= 1;
    level++;
  }
};

static uint32_t allow_mask(const Walker &w) { return 0x7u | (w.pull_possible ? 0x8u : 0u); }

// Initial state of a search (levels, visited bitmap, queue, control block),
// read back so that the root's degree is known; frontier degree sums are
// always kept, since they decide between mark and exact-once discovery.
static Walker start_search(bfsx_ctx *ctx, uint32_t root, bool pull_possible, bool stamp,
                           uint32_t *par) {
  if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
  if (root >= ctx->n_cols) fail(BFSX_ERR_ARG, "root out of range");
  hipStream_t s = ctx->stream;
  if (pull_possible) ensure_pull_state(ctx);
  if (ctx->mark_dirty) {
    BFSX_HIP(hipMemsetAsync(ctx->mark.p, 0, mark_bytes(ctx->n_cols), s));
    ctx->mark_dirty = false;
  }
  init_search(ctx, root, graph_symmetric(ctx), par);
  read_ctl(ctx);
  view_from_ctl(ctx, 0);
  Walker w;
  w.ctx = ctx;
  w.pull_possible = pull_possible;
  w.track = true;
  w.stamp = stamp;
  w.par = par;
  w.EF = ctx->view.EF;
  w.VE = ctx->view.VE;
  return w;
}

// ---------------------------------------------------------------------------
// captured level loop: K device-driven levels per hipGraph launch
// ---------------------------------------------------------------------------
void drop_loop_graph(bfsx_ctx *ctx) {
  if (ctx->loop_exec) (void)hipGraphExecDestroy(ctx->loop_exec);
  if (ctx->loop_graph) (void)hipGraphDestroy(ctx->loop_graph);
  ctx->loop_exec = nullptr;
  ctx->loop_graph = nullptr;
  ctx->loop_key = 0;
}

// fixed = BFSX_LLB / BFSX_SLB, or -1 for the scheduler's choice between them
static hipGraphExec_t loop_graph(bfsx_ctx *ctx, int fixed, bool learned) {
  const uint32_t K = (ctx->params.graph_levels + 1) & ~1u;  // even: buffers return to 0
  const MarkRule mr = mark_rule(ctx);
  const uint64_t key = 1 + ((uint64_t)(fixed + 1) | ((uint64_t)learned << 3) |
                            ((uint64_t)K << 4) | ((uint64_t)ctx->params.llb_wg << 12) |
                            ((uint64_t)ctx->params.hub_degree << 24)) ^
                       ((uint64_t)mr.lo_E * 0x9E3779B97F4A7C15ull) ^
                       ((uint64_t)mr.hi_E * 0xC2B2AE3D27D4EB4Full);
  if (ctx->loop_exec && ctx->loop_key == key) return ctx->loop_exec;
  drop_loop_graph(ctx);
  hipStream_t s = ctx->stream;
  BFSX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    for (uint32_t k = 0; k < K; k++) {
      const uint32_t cur = k & 1u;
      if (learned) level_begin(ctx);
      dev_level(ctx, cur, fixed, true, mr);
      level_end(ctx, cur, cur ^ 1, learned, fixed,
                learned ? (1u << BFSX_LLB) | (1u << BFSX_SLB) : 0u, true);
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  BFSX_HIP(hipStreamEndCapture(s, &ctx->loop_graph));
  BFSX_HIP(hipGraphInstantiate(&ctx->loop_exec, ctx->loop_graph, nullptr, nullptr, 0));
  ctx->loop_key = key;
  return ctx->loop_exec;
}

static RunOut search_graph_loop(bfsx_ctx *ctx, uint32_t root, int fixed, bool learned) {
  hipStream_t s = ctx->stream;
  start_search(ctx, root, false, learned, nullptr);
  hipGraphExec_t exec = loop_graph(ctx, learned ? -1 : (fixed < 0 ? BFSX_LLB : fixed), learned);
  BFSX_HIP(hipStreamSynchronize(s));
  ctx->trace.clear();
  RunOut out;
  BFSX_HIP(hipEventRecord(ctx->ev0, s));
  // (the root level runs LLB, as in the host loop)
  const uint64_t max_launches = (uint64_t)ctx->n_cols / 2 + 2;
  // each launch runs K level ends, each bumping the level-end counter
  const uint32_t K = (ctx->params.graph_levels + 1) & ~1u;
  uint32_t seq = ctx->h_ctl->seq;
  for (uint64_t launches = 0;; launches++) {
    if (launches > max_launches) fail(BFSX_ERR_DEVICE, "captured level loop did not terminate");
    BFSX_HIP(hipGraphLaunch(exec, s));
    seq += K;
    read_level_end(ctx, seq, 0);  // (K is even: the last level fills queue 0)
    if (ctx->view.done) break;
  }
  BFSX_HIP(hipEventRecord(ctx->ev1, s));
  BFSX_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  BFSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out.ms = ms;
  read_ctl(ctx);  // per-kind level counts
  for (int k = 0; k < BFSX_N_STEP_KINDS; k++) {
    out.steps[k] = ctx->h_ctl->kind_count[k];
    out.n_levels += out.steps[k];
  }
  return out;
}

// Beamer, Asanovic and Patterson, "Direction-Optimizing Breadth-First
// Search" (SC'12), in the form of their GAP benchmark: switch to bottom-up
// when the frontier's edges exceed the unexplored edges / alpha; stay
// bottom-up while the frontier grows or holds more than n / beta vertices.
struct BeamerRule {
  static constexpr uint64_t ALPHA = 14, BETA = 24;
  bool bottom_up = false;
  uint64_t prev_F = 0;
  int next(const bfsx_ctx *ctx, bool pull_possible, uint64_t F, uint64_t EF, uint64_t VE) {
    const uint64_t unexplored = ctx->nnz > VE ? ctx->nnz - VE : 0;
    if (!bottom_up) {
      bottom_up = pull_possible && EF > unexplored / ALPHA;
    } else if (F < prev_F && F <= (uint64_t)ctx->n_cols / BETA) {
      bottom_up = false;
    }
    prev_F = F;
    if (bottom_up) return BFSX_PULL;
    return EF * 256 >= ctx->n_cols ? BFSX_GLB : BFSX_LLB;
  }
};

static RunOut search(bfsx_ctx *ctx, uint32_t root, int strategy, uint32_t flags,
                     bool want_parents) {
  if ((flags & BFSX_RUN_DEVICE_LOOP) &&
      ((flags & (BFSX_RUN_TRACE | BFSX_RUN_STEP_TIMES | BFSX_RUN_GRAPH_LOOP)) || want_parents ||
       strategy == BFSX_BEAMER))
    fail(BFSX_ERR_UNSUPPORTED,
         "the device loop runs without traces, parents, the captured loop or the Beamer rule");
  if (strategy < BFSX_LLB || strategy > BFSX_BEAMER) fail(BFSX_ERR_ARG, "unknown strategy");
  const bool symmetric = ctx->has_graph && graph_symmetric(ctx);
  if (strategy == BFSX_PULL && !symmetric)
    fail(BFSX_ERR_UNSUPPORTED, "bottom-up steps need a symmetric graph with n_rows == n_cols");
  const bool learned = strategy == BFSX_LEARNED && ctx->sched_loaded;
  const bool beamer = strategy == BFSX_BEAMER;
  // fallback of the learned strategy without weights: fixed LLB
  const int fixed = (strategy == BFSX_LEARNED || beamer) ? BFSX_LLB : strategy;
  const bool trace = flags & BFSX_RUN_TRACE;
  const bool timed = trace || (flags & BFSX_RUN_STEP_TIMES);
  hipStream_t s = ctx->stream;
  if (flags & BFSX_RUN_GRAPH_LOOP) {
    if (beamer) fail(BFSX_ERR_UNSUPPORTED, "the direction-optimising rule runs in the host loop");
    if (fixed != BFSX_LLB && fixed != BFSX_SLB)
      fail(BFSX_ERR_UNSUPPORTED, "the captured level loop runs LLB, SLB or the learned choice between them");
    if (timed || want_parents)
      fail(BFSX_ERR_UNSUPPORTED, "the captured level loop has no per-level trace or parents");
    return search_graph_loop(ctx, root, strategy == BFSX_LEARNED ? -1 : fixed, learned);
  }
  if (want_parents) ensure(ctx->parents, sizeof(uint32_t) * ctx->n_cols, s);
  uint32_t *par = want_parents ? ctx->parents.as<uint32_t>() : nullptr;
  // initial state outside the timer (host.cpp:113-121 / run_bfs_gpu.hpp:674-677)
  Walker w = start_search(ctx, root, fixed == BFSX_PULL || ((learned || beamer) && symmetric),
                          false, par);
  BeamerRule rule;
  w.count = timed;
  BFSX_HIP(hipStreamSynchronize(s));
  ctx->trace.clear();

  RunOut out;
  BFSX_HIP(hipEventRecord(ctx->ev0, s));
  uint64_t f_edges = 0, visited_v = 1, visited_e = 0, examined = 0;
  // The root level (one vertex) runs LLB, the choice k_init leaves in the
  // control block, with no network evaluation or extra round trip; every
  // later level is picked by the network in the previous level's end.
  w.decide_allow = learned ? allow_mask(w) : 0u;
  w.decide_reuse = true;
  if (trace) {
    read_ctl(ctx);
    f_edges = ctx->h_ctl->q_edges[0];
    visited_e = ctx->h_ctl->visited_edges;
  }
  uint32_t seq = ctx->h_ctl->seq;  // level ends so far (h_ctl: the block k_init wrote)
  uint32_t decisions = ctx->h_ctl->decisions;
  const bool dev_loop = (flags & BFSX_RUN_DEVICE_LOOP) && ctx->coop_ok;
  const MarkRule mr = mark_rule(ctx);
  // speculative steps (see Walker::launch_speculative_pull / _td)
  const bool can_spec = ctx->params.speculate && learned && !timed;
  int spec = -1;  // the kind speculated for level w.level: BFSX_PULL, BFSX_LLB (LLB or SLB) or -1
  uint64_t prev_F = 1;  // the previous level's frontier size
  while (w.F > 0) {
    // a search has at most n_cols levels; more means a broken invariant
    if (w.level > ctx->n_cols) fail(BFSX_ERR_DEVICE, "level loop did not terminate");
    const int choice = learned  ? (int)ctx->view.choice
                       : beamer ? rule.next(ctx, w.pull_possible, w.F, w.EF, w.VE)
                                : fixed;
    if (dev_loop && (choice == BFSX_LLB || choice == BFSX_SLB) && w.queue_valid && !mr.marks(w.EF, w.F)) {
      // a run of thin levels on the device, up to the first heavy level, GLB
      // or PULL choice, or the end of the search
      const uint32_t level0 = w.level;
      // (at most 4096 levels per launch: the host launches again if needed)
      if (coop_levels(ctx, w.cur, learned ? -1 : fixed, w.decide_allow, learned, mr, 4096)) {
        read_level_end(ctx, ++seq, w.cur, level0);
        const uint32_t ran = (ctx->view.level - level0) & MIRROR_LEVEL_MASK;
        if (ran > 0) {
          w.cur ^= ran & 1u;
          w.level += ran;
          w.F = ctx->view.F;
          w.EF = ctx->view.EF;
          w.VE = ctx->view.VE;
          w.fbm_valid = false;
          w.queue_valid = true;
          w.last = choice;
          out.n_levels += ran;
          continue;
        }
      }
    }
    const uint32_t level = w.level, F = w.F, cur = w.cur;
    // the decision in force for this level (made at the previous level's end)
    bfsx_level_record r{};
    if (trace && learned) {
      r.sched_evaluated = ctx->h_ctl->decisions != decisions ? 1u : 0u;
      decisions = ctx->h_ctl->decisions;
      std::memcpy(r.sched_in, ctx->h_ctl->sched_in, sizeof(r.sched_in));
      std::memcpy(r.sched_out, ctx->h_ctl->sched_out, sizeof(r.sched_out));
    }
    if (timed) BFSX_HIP(hipEventRecord(ctx->ev_step0, s));
    bool new_fbm;
    if (spec == BFSX_PULL && choice == BFSX_PULL) {
      // this level's step is the speculative one already in the stream
      new_fbm = true;
      w.last = BFSX_PULL;
    } else if (spec == BFSX_LLB && (choice == BFSX_LLB || choice == BFSX_SLB) &&
               !mr.marks(w.EF, w.F)) {
      new_fbm = false;  // (the speculative top-down step writes the queue only)
      w.last = choice;
    } else {
      // (a speculative step that was not picked exits without a write)
      new_fbm = w.launch(choice);
    }
    // The guess for the next level: bottom-up after a bottom-up step, and
    // after a top-down step that wrote the next frontier's bitmap (bucketed
    // or marking: heavy, and usually followed by bottom-up levels); a thin
    // top-down step after another top-down step whose frontier did not grow
    // more than twofold (the tail of a search, a grid), not while the
    // frontier grows fast (a guess that fails costs its empty launches).
    spec = -1;
    if (can_spec && w.pull_possible && (choice == BFSX_PULL || new_fbm)) {
      w.launch_speculative_pull();
      spec = BFSX_PULL;
    } else if (can_spec && !dev_loop && choice != BFSX_PULL && !par && level > 0 &&
               (uint64_t)F <= 2ull * prev_F) {
      w.launch_speculative_td(mr);
      spec = BFSX_LLB;
    }
    prev_F = F;
    if (timed) BFSX_HIP(hipEventRecord(ctx->ev_step1, s));
    read_level_end(ctx, ++seq, cur ^ 1);
    // per-level records read the whole block (queue sizes, running totals,
    // scheduler inputs and outputs)
    if (timed) read_ctl(ctx);
    w.advance(new_fbm);
    if (timed) {
      r.level = level;
      r.choice = (uint32_t)choice;
      r.frontier_vertices = F;
      r.discovered = w.F;
      r.pull_scanned = ctx->h_ctl->examined - examined;
      examined = ctx->h_ctl->examined;
      if (trace) {
        r.frontier_edges = f_edges;
        r.unvisited_vertices = ctx->n_cols - visited_v;
        r.unvisited_edges = ctx->nnz - visited_e;
        f_edges = ctx->h_ctl->q_edges[cur ^ 1];
        visited_v = ctx->h_ctl->visited_vertices;
        visited_e = ctx->h_ctl->visited_edges;
      }
      float us = 0, kus = 0;
      BFSX_HIP(hipEventSynchronize(ctx->ev_step1));  // the mirror can arrive before it
      BFSX_HIP(hipEventElapsedTime(&us, ctx->ev_step0, ctx->ev_step1));
      BFSX_HIP(hipEventElapsedTime(&kus, ctx->ev_k0, ctx->ev_k1));
      r.us = us * 1000.0;
      r.kernel_us = kus * 1000.0;
      ctx->trace.push_back(r);
    }
    out.steps[choice]++;
    out.n_levels++;
  }
  BFSX_HIP(hipEventRecord(ctx->ev1, s));
  BFSX_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  BFSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out.ms = ms;
  if (dev_loop) {  // the device loop's levels were counted per kind on the device
    read_ctl(ctx);
    for (int k = 0; k < BFSX_N_STEP_KINDS; k++) out.steps[k] = ctx->h_ctl->kind_count[k];
  }
  return out;
}

// ---------------------------------------------------------------------------
// training rows for the scheduler: time every step kind from the same state
// ---------------------------------------------------------------------------
static void snapshot(bfsx_ctx *ctx, bool save) {
  hipStream_t s = ctx->stream;
  const size_t lb = sizeof(int32_t) * ctx->n_cols;
  const size_t vb = sizeof(uint32_t) * bitmap_words(ctx->n_cols);
  ensure(ctx->snap_levels, lb, s);
  ensure(ctx->snap_vis, vb, s);
  ensure(ctx->snap_ctl, sizeof(LevelCtl), s);
  if (save) {
    BFSX_HIP(hipMemcpyAsync(ctx->snap_levels.p, ctx->levels.p, lb, hipMemcpyDeviceToDevice, s));
    BFSX_HIP(hipMemcpyAsync(ctx->snap_vis.p, ctx->vis.p, vb, hipMemcpyDeviceToDevice, s));
    BFSX_HIP(hipMemcpyAsync(ctx->snap_ctl.p, ctx->ctl.p, sizeof(LevelCtl), hipMemcpyDeviceToDevice, s));
  } else {
    BFSX_HIP(hipMemcpyAsync(ctx->levels.p, ctx->snap_levels.p, lb, hipMemcpyDeviceToDevice, s));
    BFSX_HIP(hipMemcpyAsync(ctx->vis.p, ctx->snap_vis.p, vb, hipMemcpyDeviceToDevice, s));
    // (the level-end counter stays as it is)
    BFSX_HIP(hipMemcpyAsync(ctx->ctl.p, ctx->snap_ctl.p, CTL_RESET_BYTES, hipMemcpyDeviceToDevice, s));
  }
}

static double step_us(bfsx_ctx *ctx) {
  const LevelCtl &c = *ctx->h_ctl;
  const double ticks = c.t_end > c.t_begin ? (double)(c.t_end - c.t_begin) : 0.0;
  return ticks * 1000.0 / ctx->wall_clock_khz;
}

static uint32_t record(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots, uint32_t repeats,
                       float epsilon, uint64_t seed, float *X, float *T, float *M,
                       uint32_t max_rows) {
  if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
  if (repeats < 1) fail(BFSX_ERR_ARG, "repeats must be >= 1");
  const bool symmetric = graph_symmetric(ctx);
  // scheduler inputs come from the same kernel as at inference; the
  // network's outputs are not used here
  ensure(ctx->sched_w, sizeof(float) * BFSX_SCHED_PARAMS, ctx->stream);
  if (!ctx->sched_loaded)
    BFSX_HIP(hipMemsetAsync(ctx->sched_w.p, 0, sizeof(float) * BFSX_SCHED_PARAMS, ctx->stream));
  uint64_t rng = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next_u01 = [&]() {  // SplitMix64
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  uint32_t rows = 0;
  for (uint32_t ri = 0; ri < n_roots && rows < max_rows; ri++) {
    Walker w = start_search(ctx, roots[ri], symmetric, true, nullptr);
    const uint32_t allow = allow_mask(w);
    decide(ctx, 0, true, allow);
    w.decide_allow = allow;
    read_ctl(ctx);
    while (w.F > 0 && rows < max_rows) {
      if (w.level > ctx->n_cols) fail(BFSX_ERR_DEVICE, "level loop did not terminate");
      for (int k = 0; k < 8; k++) X[(size_t)rows * 8 + k] = ctx->h_ctl->sched_in[k];
      snapshot(ctx, true);
      const Walker saved = w;
      double best_us = 1e300;
      int best = BFSX_LLB;
      for (int kind = 0; kind < BFSX_N_STEP_KINDS; kind++) {
        const bool ok = (allow >> kind) & 1u;
        M[(size_t)rows * 4 + kind] = ok ? 1.f : 0.f;
        T[(size_t)rows * 4 + kind] = 0.f;
        if (!ok) continue;
        double t = 1e300;
        for (uint32_t rep = 0; rep < repeats; rep++) {
          snapshot(ctx, false);
          w = saved;
          // from an idle stream, as in a search, so that the time includes
          // the gaps of issuing the step's launches one after the other
          BFSX_HIP(hipStreamSynchronize(ctx->stream));
          w.launch(kind);
          read_ctl(ctx);
          t = std::min(t, step_us(ctx));
        }
        T[(size_t)rows * 4 + kind] = (float)std::log2(std::max(t, 0.01));
        if (t < best_us) {
          best_us = t;
          best = kind;
        }
      }
      int take = best;
      if (epsilon > 0 && next_u01() < epsilon) {
        int opts[4], n = 0;
        for (int kind = 0; kind < 4; kind++)
          if ((allow >> kind) & 1u) opts[n++] = kind;
        take = opts[std::min(n - 1, (int)(next_u01() * n))];
      }
      // run the chosen kind for real; the next row's inputs come from its state
      snapshot(ctx, false);
      w = saved;
      BFSX_HIP(hipMemcpyAsync(&ctx->ctl.as<LevelCtl>()->choice, &take, sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
      const bool new_fbm = w.launch(take);  // its level end decides the next level
      read_ctl(ctx);
      view_from_ctl(ctx, w.cur ^ 1);
      w.advance(new_fbm);
      rows++;
    }
  }
  BFSX_HIP(hipStreamSynchronize(ctx->stream));
  return rows;
}

static void finish_run(bfsx_ctx *ctx, const RunOut &r, int32_t *levels_out,
                       uint32_t *parents_out, bfsx_run_stats *stats) {
  hipStream_t s = ctx->stream;
  if (stats) {
    ensure(ctx->scratch, 64, s);
    reached_summary(ctx, ctx->scratch.as<unsigned long long>());
    unsigned long long h[3] = {0, 0, 0};
    BFSX_HIP(hipMemcpyAsync(h, ctx->scratch.p, sizeof(h), hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = r.ms;
    stats->edges_count = h[0];
    stats->reached = h[1];
    stats->max_level = (uint32_t)h[2];
    stats->n_levels = r.n_levels;
    for (int i = 0; i < BFSX_N_STEP_KINDS; i++) stats->steps[i] = r.steps[i];
  }
  if (levels_out)
    BFSX_HIP(hipMemcpyAsync(levels_out, ctx->levels.p, sizeof(int32_t) * ctx->n_cols,
                            hipMemcpyDeviceToHost, s));
  if (parents_out)
    BFSX_HIP(hipMemcpyAsync(parents_out, ctx->parents.p, sizeof(uint32_t) * ctx->n_cols,
                            hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
}

}  // namespace bfsx

using namespace bfsx;

// Runs `body`, converting every failure into a status code and message.
template <class F>
static int guarded(bfsx_ctx *ctx, F &&body) {
  try {
    body();
    if (ctx) ctx->err.clear();
    return BFSX_OK;
  } catch (const Error &e) {
    if (ctx) ctx->err = e.msg;
    return e.status;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->err = "host allocation failed";
    return BFSX_ERR_NOMEM;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return BFSX_ERR_DEVICE;
  } catch (...) {
    if (ctx) ctx->err = "unknown failure";
    return BFSX_ERR_DEVICE;
  }
}

extern "C" {

const char *bfsx_status_string(int st) {
  switch (st) {
    case BFSX_OK: return "ok";
    case BFSX_ERR_ARG: return "invalid argument";
    case BFSX_ERR_DEVICE: return "device error";
    case BFSX_ERR_NOMEM: return "out of memory";
    case BFSX_ERR_STATE: return "invalid state";
    case BFSX_ERR_IO: return "i/o error";
    case BFSX_ERR_OVERFLOW: return "overflow";
    case BFSX_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

const char *bfsx_last_error(const bfsx_ctx *ctx) {
  return ctx ? ctx->err.c_str() : bfsx::noalloc_last_error();
}

int bfsx_create(int device, bfsx_ctx **out) {
  if (!out) return BFSX_ERR_ARG;
  *out = nullptr;
  bfsx_ctx *ctx = nullptr;
  try {
    ctx = new bfsx_ctx;
  } catch (...) {
    return BFSX_ERR_NOMEM;
  }
  int rc = guarded(ctx, [&] {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) fail(BFSX_ERR_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) fail(BFSX_ERR_ARG, "device index out of range");
    ctx->device = device;
    BFSX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    BFSX_HIP(hipGetDeviceProperties(&prop, device));
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_per_block = (uint32_t)prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor > 0)
      ctx->lds_per_cu = (uint32_t)prop.maxSharedMemoryPerMultiProcessor;
    ctx->name = prop.gcnArchName;
    ctx->hbm_bytes = prop.totalGlobalMem;
    int coop = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) == hipSuccess)
      ctx->coop_ok = coop != 0;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
      ctx->wall_clock_khz = khz;
    BFSX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    BFSX_HIP(hipEventCreate(&ctx->ev0));
    BFSX_HIP(hipEventCreate(&ctx->ev1));
    BFSX_HIP(hipEventCreate(&ctx->ev_step0));
    BFSX_HIP(hipEventCreate(&ctx->ev_step1));
    BFSX_HIP(hipEventCreate(&ctx->ev_k0));
    BFSX_HIP(hipEventCreate(&ctx->ev_k1));
    ensure(ctx->ctl, sizeof(LevelCtl), ctx->stream);
    BFSX_HIP(hipMemsetAsync(ctx->ctl.p, 0, sizeof(LevelCtl), ctx->stream));  // level-end counter 0
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
    BFSX_HIP(hipHostMalloc((void **)&ctx->h_ctl, sizeof(LevelCtl), hipHostMallocDefault));
    BFSX_HIP(hipHostMalloc((void **)&ctx->h_mirror, sizeof(LevelMirror) * MIRROR_SLOTS,
                           hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(ctx->h_mirror, 0, sizeof(LevelMirror) * MIRROR_SLOTS);
    BFSX_HIP(hipHostGetDevicePointer((void **)&ctx->d_mirror, ctx->h_mirror, 0));
    ctx->params = default_params();
  });
  if (rc != BFSX_OK) {
    bfsx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return BFSX_OK;
}

void bfsx_destroy(bfsx_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_search_state(ctx);
  release(ctx->col_ptr);
  release(ctx->row_idx);
  release(ctx->row_idx_pull);
  drop_pull_fd(ctx);
  release(ctx->iso_bm);
  release(ctx->deg16);
  release(ctx->ctl);
  release(ctx->scratch);
  for (DevBuf *b : {&ctx->sched_w, &ctx->snap_levels, &ctx->snap_vis, &ctx->snap_ctl})
    release(*b);
  if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
  if (ctx->h_mirror) (void)hipHostFree(ctx->h_mirror);
  for (hipEvent_t e : {ctx->ev0, ctx->ev1, ctx->ev_step0, ctx->ev_step1, ctx->ev_k0, ctx->ev_k1})
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int bfsx_device_info(bfsx_ctx *ctx, char *name, size_t name_len, int *cus, uint64_t *hbm) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    if (name && name_len) {
      std::strncpy(name, ctx->name.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (cus) *cus = ctx->cu_count;
    if (hbm) *hbm = ctx->hbm_bytes;
  });
}

int bfsx_upload_csc(bfsx_ctx *ctx, uint32_t n_cols, uint32_t n_rows, const void *col_ptr,
                    int col_ptr_bytes, const uint32_t *row_idx, uint64_t nnz, uint32_t flags) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (col_ptr_bytes != 4 && col_ptr_bytes != 8) fail(BFSX_ERR_ARG, "col_ptr_bytes must be 4 or 8");
    if (n_rows > n_cols) fail(BFSX_ERR_ARG, "n_rows > n_cols");
    if (!col_ptr || (nnz && !row_idx)) fail(BFSX_ERR_ARG, "null graph pointer");
    // validate col_ptr on the host: monotone, starts at 0, ends at nnz
    const uint64_t n1 = (uint64_t)n_cols + 1;
    auto at = [&](uint64_t i) -> uint64_t {
      return col_ptr_bytes == 8 ? static_cast<const uint64_t *>(col_ptr)[i]
                                : static_cast<const uint32_t *>(col_ptr)[i];
    };
    if (at(0) != 0) fail(BFSX_ERR_ARG, "col_ptr[0] != 0");
    for (uint64_t i = 1; i < n1; i++)
      if (at(i) < at(i - 1)) fail(BFSX_ERR_ARG, "col_ptr is not monotone");
    if (at(n_rows) > nnz || at(n_cols) != nnz)
      fail(BFSX_ERR_ARG, "col_ptr end does not match nnz");
    for (uint64_t i = 0; i < nnz; i++)
      if (row_idx[i] >= n_cols) fail(BFSX_ERR_ARG, "row_idx entry >= n_cols");
    ctx->has_graph = false;
    release(ctx->col_ptr);
    release(ctx->row_idx);
    ensure(ctx->col_ptr, (size_t)col_ptr_bytes * n1, ctx->stream);
    ensure(ctx->row_idx, sizeof(uint32_t) * (nnz ? nnz : 1), ctx->stream);
    BFSX_HIP(hipMemcpyAsync(ctx->col_ptr.p, col_ptr, (size_t)col_ptr_bytes * n1,
                            hipMemcpyHostToDevice, ctx->stream));
    if (nnz)
      BFSX_HIP(hipMemcpyAsync(ctx->row_idx.p, row_idx, sizeof(uint32_t) * nnz,
                              hipMemcpyHostToDevice, ctx->stream));
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->n_cols = n_cols;
    ctx->n_rows = n_rows;
    ctx->nnz = nnz;
    ctx->ptr_bytes = col_ptr_bytes;
    ctx->gflags = flags & BFSX_GRAPH_SYMMETRIC;
    ctx->has_graph = true;
    after_graph_change(ctx);
  });
}

static uint32_t thresh32(double p) {
  const double v = p * 4294967296.0;
  if (v <= 0) return 0;
  if (v >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

int bfsx_generate_rmat(bfsx_ctx *ctx, int scale, int edge_factor, double a, double b, double c,
                       uint64_t seed_graph, uint64_t seed_perm, int col_ptr_bytes,
                       uint64_t *nnz_out) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (scale < 1 || scale > 31) fail(BFSX_ERR_ARG, "scale must be in [1,31]");
    if (edge_factor < 1 || edge_factor > 1024) fail(BFSX_ERR_ARG, "edge_factor must be in [1,1024]");
    if (col_ptr_bytes != 4 && col_ptr_bytes != 8) fail(BFSX_ERR_ARG, "col_ptr_bytes must be 4 or 8");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12) fail(BFSX_ERR_ARG, "bad quadrant probabilities");
    const uint32_t n = 1u << scale;
    const uint64_t m = (uint64_t)edge_factor << scale;
    ctx->has_graph = false;
    free_search_state(ctx);
    release(ctx->col_ptr);
    release(ctx->row_idx);
    DevBuf du, dv;
    ensure(du, sizeof(uint32_t) * m, ctx->stream);
    ensure(dv, sizeof(uint32_t) * m, ctx->stream);
    gen_rmat_pairs(ctx, scale, m, thresh32(a), thresh32(a + b), thresh32(a + b + c), seed_graph,
                   seed_perm, du.as<uint32_t>(), dv.as<uint32_t>());
    build_symmetric_csc(ctx, n, du.as<uint32_t>(), dv.as<uint32_t>(), m, col_ptr_bytes);
    release(du);
    release(dv);
    after_graph_change(ctx);
    if (nnz_out) *nnz_out = ctx->nnz;
  });
}

int bfsx_generate_grid(bfsx_ctx *ctx, uint32_t side, double drop_prob, uint32_t n_long,
                       uint64_t seed, int col_ptr_bytes, uint64_t *nnz_out) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (side < 2 || side > 65535) fail(BFSX_ERR_ARG, "side must be in [2,65535]");
    if (drop_prob < 0 || drop_prob > 1) fail(BFSX_ERR_ARG, "drop_prob must be in [0,1]");
    if (col_ptr_bytes != 4 && col_ptr_bytes != 8) fail(BFSX_ERR_ARG, "col_ptr_bytes must be 4 or 8");
    const uint64_t V = (uint64_t)side * side;
    const uint64_t m = 2ull * side * (side - 1) + n_long;
    ctx->has_graph = false;
    free_search_state(ctx);
    release(ctx->col_ptr);
    release(ctx->row_idx);
    DevBuf du, dv;
    ensure(du, sizeof(uint32_t) * m, ctx->stream);
    ensure(dv, sizeof(uint32_t) * m, ctx->stream);
    gen_grid_pairs(ctx, side, thresh32(drop_prob), n_long, seed, du.as<uint32_t>(), dv.as<uint32_t>());
    build_symmetric_csc(ctx, (uint32_t)V, du.as<uint32_t>(), dv.as<uint32_t>(), m, col_ptr_bytes);
    release(du);
    release(dv);
    after_graph_change(ctx);
    if (nnz_out) *nnz_out = ctx->nnz;
  });
}

int bfsx_graph_info_get(bfsx_ctx *ctx, bfsx_graph_info *out) {
  if (!ctx || !out) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    out->n_rows = ctx->n_rows;
    out->n_cols = ctx->n_cols;
    out->nnz = ctx->nnz;
    out->col_ptr_bytes = ctx->ptr_bytes;
    out->flags = ctx->gflags;
    out->max_degree = ctx->max_degree;
    out->n_isolated = ctx->n_isolated;
    out->pull_hubs = ctx->pull_hubs;
    out->hot_vertices = ctx->hot_n;
  });
}

int bfsx_order_adjacency(bfsx_ctx *ctx, int order) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    if (order != BFSX_ORDER_ID && order != BFSX_ORDER_HUB_FIRST &&
        order != BFSX_ORDER_PULL_HUB_FIRST)
      fail(BFSX_ERR_ARG,
           "order must be BFSX_ORDER_ID, BFSX_ORDER_HUB_FIRST or BFSX_ORDER_PULL_HUB_FIRST");
    if (order == BFSX_ORDER_PULL_HUB_FIRST) {
      ensure(ctx->row_idx_pull, sizeof(uint32_t) * std::max<uint64_t>(ctx->nnz, 1), ctx->stream);
      order_adjacency(ctx, true, ctx->row_idx_pull.as<uint32_t>());
      if (ctx->max_degree < 0xFFFFFFFFull) {  // degrees fit the array's 32 bits
        ensure(ctx->pull_fd, sizeof(uint2) * std::max<uint64_t>(ctx->n_rows, 1), ctx->stream);
        first_neighbours(ctx, ctx->row_idx_pull.as<uint32_t>(), ctx->pull_fd.as<uint2>());
        index_hub_neighbours(ctx);
        compact_first_neighbours(ctx);
      }
      build_hot_copy(ctx);
      ctx->gflags |= BFSX_GRAPH_PULL_COPY;
      return;
    }
    release(ctx->row_idx_pull);
    drop_pull_fd(ctx);
    ctx->gflags &= ~BFSX_GRAPH_PULL_COPY;
    order_adjacency(ctx, order == BFSX_ORDER_HUB_FIRST, ctx->row_idx.as<uint32_t>());
    if (order == BFSX_ORDER_HUB_FIRST) ctx->gflags |= BFSX_GRAPH_HUB_FIRST;
    else ctx->gflags &= ~BFSX_GRAPH_HUB_FIRST;
  });
}

int bfsx_download_csc(bfsx_ctx *ctx, void *col_ptr, uint32_t *row_idx) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    if (col_ptr)
      BFSX_HIP(hipMemcpyAsync(col_ptr, ctx->col_ptr.p,
                              (size_t)ctx->ptr_bytes * ((uint64_t)ctx->n_cols + 1),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (row_idx && ctx->nnz)
      BFSX_HIP(hipMemcpyAsync(row_idx, ctx->row_idx.p, sizeof(uint32_t) * ctx->nnz,
                              hipMemcpyDeviceToHost, ctx->stream));
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int bfsx_pick_roots(bfsx_ctx *ctx, uint32_t n_roots, uint64_t seed, int mode, uint32_t *roots_out,
                    uint32_t *n_found) {
  if (!ctx || (n_roots && !roots_out)) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    if (mode < 0 || mode > 2)
      fail(BFSX_ERR_ARG, "mode must be 0 (degree >= 1), 1 (largest component) or 2 (top degree)");
    hipStream_t s = ctx->stream;
    if (mode == 2) {
      const uint32_t k = std::min(n_roots, ctx->n_rows);
      DevBuf dv;
      ensure(dv, sizeof(uint32_t) * std::max(k, 1u), s);
      top_degree_vertices(ctx, k, dv.as<uint32_t>());
      if (k) {
        BFSX_HIP(hipMemcpyAsync(roots_out, dv.p, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s));
        BFSX_HIP(hipStreamSynchronize(s));
      }
      release(dv);
      if (n_found) *n_found = k;
      return;
    }
    const int32_t *comp = nullptr;
    std::vector<int32_t> keep_levels;
    if (mode == 1) {
      if (!(ctx->gflags & BFSX_GRAPH_SYMMETRIC)) fail(BFSX_ERR_UNSUPPORTED, "components need a symmetric graph");
      // A component holding more than half the vertices is the largest one.
      // Search from degree>=1 candidates until one reaches more than V/2.
      const uint32_t probe_batch = 64;
      std::vector<uint32_t> cand(probe_batch);
      DevBuf dv, dok;
      ensure(dv, sizeof(uint32_t) * probe_batch, s);
      ensure(dok, probe_batch, s);
      std::vector<uint8_t> ok(probe_batch);
      bool found = false;
      for (uint64_t first = 0; first < 64ull * probe_batch && !found; first += probe_batch) {
        root_candidates(ctx, seed ^ 0x5bd1e995ull, first, probe_batch, nullptr, dv.as<uint32_t>(),
                        dok.as<uint8_t>());
        BFSX_HIP(hipMemcpyAsync(cand.data(), dv.p, sizeof(uint32_t) * probe_batch, hipMemcpyDeviceToHost, s));
        BFSX_HIP(hipMemcpyAsync(ok.data(), dok.p, probe_batch, hipMemcpyDeviceToHost, s));
        BFSX_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < probe_batch && !found; i++) {
          if (!ok[i]) continue;
          search(ctx, cand[i], BFSX_LLB, 0, false);
          bfsx_run_stats st;
          finish_run(ctx, RunOut{}, nullptr, nullptr, &st);
          if (st.reached * 2 > ctx->n_cols) found = true;
        }
      }
      release(dv);
      release(dok);
      if (!found) fail(BFSX_ERR_UNSUPPORTED, "no component holds more than half the vertices");
      comp = ctx->levels.as<int32_t>();
    }
    const uint32_t batch = 1u << 16;
    DevBuf dv, dok;
    ensure(dv, sizeof(uint32_t) * batch, s);
    ensure(dok, batch, s);
    std::vector<uint32_t> cand(batch);
    std::vector<uint8_t> ok(batch);
    std::vector<uint32_t> got;
    const uint64_t max_tries = 64ull * ctx->n_cols + 1024;
    for (uint64_t first = 0; first < max_tries && got.size() < n_roots; first += batch) {
      root_candidates(ctx, seed, first, batch, comp, dv.as<uint32_t>(), dok.as<uint8_t>());
      BFSX_HIP(hipMemcpyAsync(cand.data(), dv.p, sizeof(uint32_t) * batch, hipMemcpyDeviceToHost, s));
      BFSX_HIP(hipMemcpyAsync(ok.data(), dok.p, batch, hipMemcpyDeviceToHost, s));
      BFSX_HIP(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < batch && got.size() < n_roots && first + i < max_tries; i++) {
        if (!ok[i]) continue;
        if (std::find(got.begin(), got.end(), cand[i]) == got.end()) got.push_back(cand[i]);
      }
    }
    release(dv);
    release(dok);
    std::copy(got.begin(), got.end(), roots_out);
    if (n_found) *n_found = (uint32_t)got.size();
  });
}

int bfsx_run(bfsx_ctx *ctx, uint32_t root, int strategy, uint32_t run_flags, int32_t *levels_out,
             uint32_t *parents_out, bfsx_run_stats *stats) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    const RunOut r = search(ctx, root, strategy, run_flags, parents_out != nullptr);
    finish_run(ctx, r, levels_out, parents_out, stats);
  });
}

int bfsx_run_batch(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots, int strategy,
                   uint32_t run_flags, int32_t *levels_out, bfsx_run_stats *stats) {
  if (!ctx || (n_roots && !roots)) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    for (uint32_t i = 0; i < n_roots; i++) {
      const RunOut r = search(ctx, roots[i], strategy, run_flags, false);
      finish_run(ctx, r, levels_out ? levels_out + (size_t)i * ctx->n_cols : nullptr, nullptr,
                 stats ? stats + i : nullptr);
    }
  });
}

int bfsx_last_trace(bfsx_ctx *ctx, bfsx_level_record *out, uint32_t max_records, uint32_t *n) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    const uint32_t k = (uint32_t)std::min<size_t>(ctx->trace.size(), max_records);
    if (out && k) std::copy(ctx->trace.begin(), ctx->trace.begin() + k, out);
    if (n) *n = (uint32_t)ctx->trace.size();
  });
}

int bfsx_level_histogram(bfsx_ctx *ctx, uint64_t *counts, uint64_t *edges, uint32_t max_levels,
                         uint32_t *n_levels) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph || !ctx->levels.p) fail(BFSX_ERR_STATE, "no search has run");
    if (max_levels == 0 || max_levels > (1u << 24)) fail(BFSX_ERR_ARG, "max_levels out of range");
    hipStream_t s = ctx->stream;
    ensure(ctx->scratch, sizeof(unsigned long long) * (2 * (size_t)max_levels + 3), s);
    unsigned long long *d = ctx->scratch.as<unsigned long long>();
    reached_summary(ctx, d);
    level_histogram(ctx, max_levels, d + 3, d + 3 + max_levels);
    std::vector<unsigned long long> h(2 * (size_t)max_levels + 3);
    BFSX_HIP(hipMemcpyAsync(h.data(), d, sizeof(unsigned long long) * h.size(),
                            hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    const uint32_t nl = h[1] ? (uint32_t)h[2] + 1 : 0;
    if (n_levels) *n_levels = nl;
    if (nl > max_levels) fail(BFSX_ERR_OVERFLOW, "more levels than max_levels");
    for (uint32_t i = 0; i < max_levels; i++) {
      if (counts) counts[i] = h[3 + i];
      if (edges) edges[i] = h[3 + max_levels + i];
    }
  });
}

int bfsx_stream_copy(bfsx_ctx *ctx, uint64_t bytes, int repeats, double *ms) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (bytes < 16 || repeats < 1) fail(BFSX_ERR_ARG, "bytes >= 16 and repeats >= 1 required");
    bytes &= ~15ull;
    DevBuf a, b;
    ensure(a, bytes, ctx->stream);
    ensure(b, bytes, ctx->stream);
    BFSX_HIP(hipMemsetAsync(a.p, 1, bytes, ctx->stream));
    stream_copy(ctx, a.p, b.p, bytes);  // warm-up
    BFSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < repeats; i++) stream_copy(ctx, a.p, b.p, bytes);
    BFSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    BFSX_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0;
    BFSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t / repeats;
    release(a);
    release(b);
  });
}

int bfsx_run_multi(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits,
                   uint32_t n_sources, uint32_t n_bits, uint8_t *levels_out,
                   bfsx_batch_stats *stats) {
  if (!ctx || (n_sources && (!sources || !bits))) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    // the queue buffers and control block are shared with single searches
    const MsResult r = ms_run(ctx, sources, bits, n_sources, n_bits, levels_out);
    if (stats) {
      stats->ms = r.ms;
      stats->edges_count = r.edges;
      stats->n_levels = r.n_levels;
      stats->push_levels = r.push_levels;
      stats->pull_levels = r.pull_levels;
    }
  });
}

int bfsx_generate_rmat_union(bfsx_ctx *ctx, uint32_t n_graphs, const int *scales,
                             const int *edge_factors, const uint64_t *seeds_graph,
                             const uint64_t *seeds_perm, int col_ptr_bytes, uint32_t *offsets,
                             uint64_t *nnz_out) {
  if (!ctx || n_graphs == 0 || !scales || !edge_factors || !seeds_graph || !seeds_perm)
    return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (col_ptr_bytes != 4 && col_ptr_bytes != 8) fail(BFSX_ERR_ARG, "col_ptr_bytes must be 4 or 8");
    uint64_t V = 0, M = 0;
    std::vector<uint64_t> off(n_graphs + 1, 0), eoff(n_graphs + 1, 0);
    for (uint32_t g = 0; g < n_graphs; g++) {
      if (scales[g] < 1 || scales[g] > 30) fail(BFSX_ERR_ARG, "scale must be in [1,30]");
      if (edge_factors[g] < 1 || edge_factors[g] > 1024) fail(BFSX_ERR_ARG, "bad edge factor");
      off[g] = V;
      eoff[g] = M;
      V += 1ull << scales[g];
      M += (uint64_t)edge_factors[g] << scales[g];
    }
    off[n_graphs] = V;
    eoff[n_graphs] = M;
    if (V >= 0xFFFFFFFFull) fail(BFSX_ERR_ARG, "packed graphs exceed 2^32-1 vertices");
    ctx->has_graph = false;
    free_search_state(ctx);
    release(ctx->col_ptr);
    release(ctx->row_idx);
    DevBuf du, dv;
    ensure(du, sizeof(uint32_t) * M, ctx->stream);
    ensure(dv, sizeof(uint32_t) * M, ctx->stream);
    for (uint32_t g = 0; g < n_graphs; g++)
      gen_rmat_pairs_offset(ctx, scales[g], eoff[g + 1] - eoff[g], thresh32(0.57),
                            thresh32(0.57 + 0.19), thresh32(0.57 + 0.19 + 0.19),
                            seeds_graph[g], seeds_perm[g],
                            (uint32_t)off[g], du.as<uint32_t>() + eoff[g],
                            dv.as<uint32_t>() + eoff[g]);
    build_symmetric_csc(ctx, (uint32_t)V, du.as<uint32_t>(), dv.as<uint32_t>(), M, col_ptr_bytes);
    release(du);
    release(dv);
    after_graph_change(ctx);
    if (offsets)
      for (uint32_t g = 0; g <= n_graphs; g++) offsets[g] = (uint32_t)off[g];
    if (nnz_out) *nnz_out = ctx->nnz;
  });
}

int bfsx_sched_set_weights(bfsx_ctx *ctx, const float *w, uint32_t n) {
  if (!ctx) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!w) {
      ctx->sched_loaded = false;
      return;
    }
    if (n != BFSX_SCHED_PARAMS) fail(BFSX_ERR_ARG, "scheduler needs 1476 parameters");
    for (uint32_t i = 0; i < n; i++)
      if (!std::isfinite(w[i])) fail(BFSX_ERR_ARG, "scheduler weight is not finite");
    ensure(ctx->sched_w, sizeof(float) * n, ctx->stream);
    BFSX_HIP(hipMemcpyAsync(ctx->sched_w.p, w, sizeof(float) * n, hipMemcpyHostToDevice,
                            ctx->stream));
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->sched_loaded = true;
  });
}

int bfsx_mlp_forward(bfsx_ctx *ctx, const float *X, uint32_t n, const float *W, float *Y) {
  if (!ctx || !X || !W || !Y || n == 0) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    mlp_forward(ctx, X, n, W, Y);
  });
}

int bfsx_mlp_loss_grad(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
                       const float *W, float *grad, float *loss) {
  if (!ctx || !X || !T || !M || !W || !grad || !loss || n == 0) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    mlp_loss_grad(ctx, X, T, M, n, W, grad, loss);
  });
}

int bfsx_mlp_train(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
                   float *W, uint32_t epochs, float lr, float *loss_curve) {
  if (!ctx || !X || !T || !M || !W || n == 0) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!(lr > 0) || epochs > 1000000) fail(BFSX_ERR_ARG, "bad learning rate or epoch count");
    mlp_train(ctx, X, T, M, n, W, epochs, lr, loss_curve);
  });
}

int bfsx_sched_record(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots, uint32_t repeats,
                      float epsilon, uint64_t seed, float *X, float *T, float *M,
                      uint32_t max_rows, uint32_t *n_rows) {
  if (!ctx || (n_roots && !roots) || !X || !T || !M) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    const uint32_t r = record(ctx, roots, n_roots, repeats, epsilon, seed, X, T, M, max_rows);
    if (n_rows) *n_rows = r;
  });
}

// Harmonic-mean GTEPS of searches from `roots` with the context's parameters.
static double tune_score(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n, int strategy,
                         const std::vector<double> &edges) {
  double s = 0;
  for (uint32_t i = 0; i < n; i++) {
    const RunOut r = search(ctx, roots[i], strategy, 0, false);
    s += r.ms * 1e-3 / std::max(edges[i], 1.0);
  }
  return s > 0 ? n / s / 1e9 : 0.0;
}

int bfsx_autotune(bfsx_ctx *ctx, const uint32_t *roots, uint32_t n_roots, int strategy,
                  uint32_t passes, bfsx_params *best, double *gteps_before, double *gteps_after) {
  if (!ctx || !roots || n_roots == 0) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    if (strategy < BFSX_LLB || strategy > BFSX_BEAMER) fail(BFSX_ERR_ARG, "unknown strategy");
    if (passes < 1 || passes > 4) fail(BFSX_ERR_ARG, "passes must be in [1,4]");
    std::vector<double> edges(n_roots);
    for (uint32_t i = 0; i < n_roots; i++) {
      const RunOut r = search(ctx, roots[i], strategy, 0, false);
      bfsx_run_stats st;
      finish_run(ctx, r, nullptr, nullptr, &st);
      edges[i] = (double)st.edges_count;
    }
    struct Knob {
      uint32_t bfsx_params::*field;
      std::vector<uint32_t> values;
      uint32_t kinds;  // step kinds whose kernels read this parameter
    };
    const uint32_t L = 1u << BFSX_LLB, G = 1u << BFSX_GLB, S = 1u << BFSX_SLB, P = 1u << BFSX_PULL;
    const std::vector<Knob> knobs = {
        {&bfsx_params::llb_wg, {64, 128, 256, 512, 1024}, L | S},
        {&bfsx_params::hub_degree, {0, 256, 1024, 4096, 16384}, L | S},
        {&bfsx_params::mark_ratio_q10, {4, 16, 64, 256, 1024}, L | S | G},
        {&bfsx_params::bucket_ratio_q10, {0, 256, 512, 768, 0xFFFFFFFFu}, G},
        {&bfsx_params::glb_wg, {128, 256, 512, 1024}, G},
        {&bfsx_params::glb_blocks_per_cu, {2, 4, 8, 16, 32}, G},
        {&bfsx_params::glb_chunk, {1024, 2048, 4096, 8192, 16384}, G},
        {&bfsx_params::pull_wg, {64, 128, 256, 512}, P},
    };
    const uint32_t used = strategy == BFSX_LEARNED  ? (L | G | S | P)
                          : strategy == BFSX_BEAMER ? (L | G | P)
                                                    : (1u << strategy);
    bfsx_params best_p = ctx->params;
    double best_s = tune_score(ctx, roots, n_roots, strategy, edges);
    const double start_s = best_s;
    for (uint32_t pass = 0; pass < passes; pass++) {
      for (const Knob &k : knobs) {
        if (!(k.kinds & used)) continue;
        for (uint32_t v : k.values) {
          if (best_p.*(k.field) == v) continue;
          bfsx_params p = best_p;
          p.*(k.field) = v;
          validate_params(p);
          ctx->params = p;
          drop_loop_graph(ctx);
          const double sc = tune_score(ctx, roots, n_roots, strategy, edges);
          if (sc > best_s * 1.01) {  // ignore differences within timing noise
            best_s = sc;
            best_p = p;
          }
        }
      }
    }
    ctx->params = best_p;
    drop_loop_graph(ctx);
    if (best) *best = best_p;
    if (gteps_before) *gteps_before = start_s;
    if (gteps_after) *gteps_after = best_s;
  });
}

int bfsx_autotune_multi(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits,
                        uint32_t n_sources, uint32_t n_bits, uint32_t passes, bfsx_params *best,
                        double *gteps_before, double *gteps_after) {
  if (!ctx || !sources || !bits || n_sources == 0) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    set_device(ctx);
    if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
    if (passes < 1 || passes > 4) fail(BFSX_ERR_ARG, "passes must be in [1,4]");
    struct Knob {
      uint32_t bfsx_params::*field;
      std::vector<uint32_t> values;
    };
    const std::vector<Knob> knobs = {
        {&bfsx_params::ms_relabel, {0, 1}},
        {&bfsx_params::ms_push_divisor, {4, 8, 14, 20, 32}},
        {&bfsx_params::ms_pull_hub, {64, 128, 256, 512, 1024}},
        {&bfsx_params::ms_graph_levels, {2, 4, 8, 16}},
    };
    // aggregate throughput of the batch (best of two runs)
    auto score = [&]() {
      double t = 1e300;
      unsigned long long e = 0;
      for (int rep = 0; rep < 2; rep++) {
        const MsResult r = ms_run(ctx, sources, bits, n_sources, n_bits, nullptr);
        t = std::min(t, r.ms);
        e = r.edges;
      }
      return t > 0 ? (double)e / (t * 1e-3) / 1e9 : 0.0;
    };
    bfsx_params best_p = ctx->params;
    double best_s = score();
    const double start_s = best_s;
    for (uint32_t pass = 0; pass < passes; pass++) {
      for (const Knob &k : knobs) {
        for (uint32_t v : k.values) {
          if (best_p.*(k.field) == v) continue;
          bfsx_params p = best_p;
          p.*(k.field) = v;
          validate_params(p);
          ctx->params = p;
          const double sc = score();
          if (sc > best_s * 1.01) {
            best_s = sc;
            best_p = p;
          }
        }
      }
    }
    ctx->params = best_p;
    if (best) *best = best_p;
    if (gteps_before) *gteps_before = start_s;
    if (gteps_after) *gteps_after = best_s;
  });
}

int bfsx_get_params(bfsx_ctx *ctx, bfsx_params *out) {
  if (!ctx || !out) return BFSX_ERR_ARG;
  *out = ctx->params;
  return BFSX_OK;
}

int bfsx_set_params(bfsx_ctx *ctx, const bfsx_params *in) {
  if (!ctx || !in) return BFSX_ERR_ARG;
  return guarded(ctx, [&] {
    validate_params(*in);
    ctx->params = *in;
    drop_loop_graph(ctx);
  });
}

}  // extern "C"
