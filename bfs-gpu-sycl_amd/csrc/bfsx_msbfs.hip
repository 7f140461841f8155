// Batched multi-source BFS: up to 256 searches advance together, one bit per
// search in a W-word (W = 1, 2, 4 or 8) bitset per vertex.
//
//   seen[u]  bits of the searches that have reached u
//   cur[u]   bits that reached u at the current level (the frontier)
//   nxt[u]   bits that reach u at the next level
//   lv[u]    uint8 level per search, vertex-major (32*W bytes per vertex),
//            transposed to one row per search at the end
//
// A level is either top-down (few active vertices: each active vertex ORs its
// frontier words into its neighbours' nxt words with atomicOr, then a
// finalize pass keeps only unseen bits) or bottom-up (many: a group of W
// lanes owns a vertex and ORs its neighbours' frontier words, word j in lane
// j, stopping as soon as every search has reached the vertex).  Both write
// levels for the new bits and append vertices with new bits to the active
// queue with one atomic per wave.  A source set may hold several vertices
// for one bit; the result is then the distance to the nearest of them, which
// is how disjoint graphs packed into one CSC run in the same launch.
#include <algorithm>

#include "bfsx_internal.hpp"

namespace bfsx {

namespace {

__device__ __forceinline__ uint32_t ms_lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Appends u once per group (the group's first lane) with one atomic per wave.
__device__ __forceinline__ void ms_append(bool leader_with_new, uint32_t u, uint32_t *q,
                                          uint32_t *qc) {
  const unsigned long long m = __ballot(leader_with_new);
  if (!m) return;
  const int first = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == first) base = atomicAdd(qc, (uint32_t)__popcll(m));
  base = __shfl(base, first, 64);
  if (leader_with_new) q[base + ms_lanes_below(m)] = u;
}

// Writes level L for the set bits of `bits` (word j of vertex u).
__device__ __forceinline__ void ms_write_levels(uint8_t *lv, uint32_t NB, uint32_t u, uint32_t j,
                                                uint32_t bits, uint8_t L) {
  uint8_t *p = lv + (size_t)u * NB + 32 * j;
  for (uint32_t b = bits; b; b &= b - 1) p[__ffs(b) - 1] = L;
}

}  // namespace

struct MsArgs {
  const void *col_ptr;
  const uint32_t *row_idx;
  uint32_t n_rows, n_cols;
  uint32_t *seen, *cur, *nxt;
  uint8_t *lv;
  uint32_t *q_in, *q_out, *q_count_out;
  uint32_t F;
  uint32_t level;  // level being written (the new bits' level)
  uint32_t *overflow;
  unsigned long long *q_edges_out;  // out-degree sum of the next active queue
  // one bit per vertex: in some search's frontier (cur != 0); act_cur is
  // null on the first level.  Bottom-up levels probe it (a few MiB, L2
  // resident) before reading a neighbour's bitset, whose 32 bytes cost a
  // whole line.
  const uint32_t *act_cur;
  uint32_t *act_next;
};

// Sets the act_next bits of the G = 64 / W consecutive vertices of a wave
// (first one u0) whose group found new bits (m: ballot of nb != 0).
template <int W>
__device__ __forceinline__ void ms_mark_active(uint32_t *act_next, uint32_t u0,
                                               unsigned long long m) {
  constexpr uint32_t G = 64 / W;
  if ((threadIdx.x & 63) != 0 || !m) return;
  unsigned long long bits = 0;
#pragma unroll
  for (uint32_t g = 0; g < G; g++)
    if ((m >> (g * W)) & ((W == 64) ? ~0ull : ((1ull << W) - 1))) bits |= 1ull << g;
  // u0 is a multiple of G, so the G bits stay inside one or two words
  const uint32_t w = u0 >> 5, sh = u0 & 31;
  const unsigned long long x = bits << sh;
  if ((uint32_t)x) atomicOr(&act_next[w], (uint32_t)x);
  if ((uint32_t)(x >> 32)) atomicOr(&act_next[w + 1], (uint32_t)(x >> 32));
}


// Word j of the frontier bitsets of eight neighbours, ORed.  The W-lane group
// starting at lane g0 holds the ids, PER = 8 / W per lane (0xFFFFFFFF past
// the list's end).  Every load is unconditional (a skipped neighbour reads
// vertex 0's hot line instead), so each phase has its eight loads in flight
// together: the active-map words, then the bitset words.
template <int W>
__device__ __forceinline__ uint32_t ms_gather8(const MsArgs &a, const uint32_t (&ids)[8 / W],
                                               uint32_t g0, uint32_t j) {
  uint32_t w[8];
  bool ok[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    w[i] = __shfl(ids[i / W], (int)(g0 + (i % W)), 64);
    ok[i] = w[i] != 0xFFFFFFFFu;
    w[i] = ok[i] ? w[i] : 0u;
  }
  if (a.act_cur) {
    uint32_t aw[8];
#pragma unroll
    for (int i = 0; i < 8; i++) aw[i] = a.act_cur[w[i] >> 5];
#pragma unroll
    for (int i = 0; i < 8; i++) ok[i] = ok[i] && ((aw[i] >> (w[i] & 31)) & 1u);
  }
  uint32_t x[8];
#pragma unroll
  for (int i = 0; i < 8; i++) x[i] = a.cur[(size_t)(ok[i] ? w[i] : 0u) * W + j];
  uint32_t part = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) part |= ok[i] ? x[i] : 0u;
  return part;
}

// Adds deg(u) of each group leader that appended u, one atomic per wave.
template <typename PtrT>
__device__ __forceinline__ void ms_track(const PtrT *col_ptr, uint32_t n_rows, bool appended,
                                         uint32_t u, unsigned long long *acc) {
  unsigned long long d = (appended && u < n_rows) ? (unsigned long long)(col_ptr[u + 1] - col_ptr[u]) : 0ull;
  for (int k = 32; k > 0; k >>= 1) d += __shfl_xor(d, k, 64);
  if ((threadIdx.x & 63) == 0 && d) atomicAdd(acc, d);
}

// ---- top-down: expand active vertices, atomicOr frontier words into nxt ----
// Active vertices with more than MS_HUB edges go to a list that every wave
// then splits (k_ms_push_hubs); the rest are expanded by one wave each.
constexpr uint32_t MS_HUB = 2048;
// Bottom-up levels hand vertices with more neighbours than this to
// k_ms_pull_hubs (a work-group each): a W-lane group scanning a long list
// alone would hold its wave for hundreds of rounds.
constexpr uint32_t MS_PULL_HUB = 256;

template <typename PtrT, int W>
__device__ __forceinline__ void ms_push_edges(const MsArgs &a, const PtrT *col_ptr, uint32_t v,
                                              PtrT first, PtrT stride) {
  const uint32_t j = (threadIdx.x & 63) % W;
  const uint32_t f = a.cur[(size_t)v * W + j];
  const PtrT e = col_ptr[v + 1];
  for (PtrT k = col_ptr[v] + first; k < e; k += stride) {
    const uint32_t u = a.row_idx[k];
    if (f) {
      const uint32_t nb = f & ~a.seen[(size_t)u * W + j];
      if (nb) atomicOr(&a.nxt[(size_t)u * W + j], nb);
    }
  }
}

template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_push(MsArgs a, uint32_t *hubs, uint32_t *n_hubs) {
  // a group of W lanes per edge: lane j handles word j
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t G = 64 / W;  // groups per wave
  const uint32_t lane = threadIdx.x & 63, grp = lane / W;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t i = wave; i < a.F; i += nwaves) {
    const uint32_t v = a.q_in[i];
    if (v >= a.n_rows) continue;
    if ((uint64_t)(col_ptr[v + 1] - col_ptr[v]) > MS_HUB) {
      if (lane == 0) hubs[atomicAdd(n_hubs, 1u)] = v;
      continue;
    }
    ms_push_edges<PtrT, W>(a, col_ptr, v, (PtrT)grp, (PtrT)G);
  }
}

template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_push_hubs(MsArgs a, const uint32_t *hubs,
                                                      const uint32_t *n_hubs) {
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t G = 64 / W;
  const uint32_t grp = (threadIdx.x & 63) / W;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * 4;
  const uint32_t nh = *n_hubs;
  for (uint32_t h = 0; h < nh; h++)
    ms_push_edges<PtrT, W>(a, col_ptr, hubs[h], (PtrT)(wave * G + grp), (PtrT)nwaves * G);
}

// ---- after a top-down level: keep unseen bits, levels, next active queue ----
template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_finalize(MsArgs a, uint32_t n_groups_total) {
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t G = 64 / W;
  const uint32_t lane = threadIdx.x & 63, j = lane % W, grp = lane / W;
  const uint32_t NB = 32 * W;
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6); wv < n_groups_total; wv += nwaves) {
    const uint32_t u = wv * G + grp;
    uint32_t nb = 0;
    if (u < a.n_cols) {
      const size_t o = (size_t)u * W + j;
      const uint32_t x = a.nxt[o];
      if (x) {
        const uint32_t s = a.seen[o];
        nb = x & ~s;
        if (nb) a.seen[o] = s | nb;
        a.nxt[o] = nb;
      }
    }
    if (nb) {
      if (a.level > 254) atomicOr(a.overflow, 1u);
      else ms_write_levels(a.lv, NB, u, j, nb, (uint8_t)a.level);
    }
    const unsigned long long m = __ballot(nb != 0);
    const bool group_new = (m >> (grp * W)) & ((W == 64) ? ~0ull : ((1ull << W) - 1));
    const bool app = group_new && j == 0 && u < a.n_cols;
    ms_append(app, u, a.q_out, a.q_count_out);
    ms_track<PtrT>(col_ptr, a.n_rows, app, u, a.q_edges_out);
    ms_mark_active<W>(a.act_next, wv * G, m);
  }
}

// ---- bottom-up: every vertex with unseen bits ORs its neighbours' frontier ----
template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_pull(MsArgs a, uint32_t n_groups_total) {
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t G = 64 / W;
  const uint32_t lane = threadIdx.x & 63, j = lane % W, grp = lane / W;
  const uint32_t NB = 32 * W;
  const uint32_t nwaves = gridDim.x * 4;
  const unsigned long long gmask = ((W == 64) ? ~0ull : ((1ull << W) - 1)) << (grp * W);
  for (uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6); wv < n_groups_total; wv += nwaves) {
    const uint32_t u = wv * G + grp;
    const bool valid = u < a.n_cols;
    const size_t o = (size_t)u * W + j;
    const uint32_t s = valid ? a.seen[o] : ~0u;
    // a group whose vertex every search has reached (or that cannot expand)
    // skips; vertices with more than MS_PULL_HUB neighbours are k_ms_pull_hubs'
    const bool row = valid && u < a.n_rows;
    const PtrT b = row ? col_ptr[u] : PtrT(0), e = row ? col_ptr[u + 1] : PtrT(0);
    const bool hub = (uint64_t)(e - b) > MS_PULL_HUB;
    bool todo = row && !hub && s != ~0u;
    const bool group_todo = (__ballot(todo) & gmask) != 0;
    uint32_t acc = 0;
    if (group_todo) {
      // eight neighbours per round: the group's lanes load their ids in one
      // coalesced read, then every lane loads its word of all eight frontier
      // bitsets (eight independent loads in flight)
      constexpr int PER = 8 / W;
      for (PtrT k = b; k < e; k += 8) {
        uint32_t ids[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) {
          const PtrT kk = k + (PtrT)(t * W + j);
          const uint32_t r = a.row_idx[kk < e ? kk : b];  // b < e: no branch around the load
          ids[t] = kk < e ? r : 0xFFFFFFFFu;
        }
        acc |= ms_gather8<W>(a, ids, grp * W, j);
        // stop once every unseen bit of the group's vertex has arrived
        const bool full = (acc | s) == ~0u;
        if ((__ballot(!full) & gmask) == 0) break;
      }
    }
    const uint32_t nb = acc & ~s;
    if (valid && !hub) {
      if (nb) a.seen[o] = s | nb;
      a.nxt[o] = nb;
    }
    if (nb) {
      if (a.level > 254) atomicOr(a.overflow, 1u);
      else ms_write_levels(a.lv, NB, u, j, nb, (uint8_t)a.level);
    }
    const unsigned long long m = __ballot(nb != 0);
    const bool app = (m & gmask) != 0 && j == 0 && valid;
    ms_append(app, u, a.q_out, a.q_count_out);
    ms_track<PtrT>(col_ptr, a.n_rows, app, u, a.q_edges_out);
    ms_mark_active<W>(a.act_next, wv * G, m);
  }
}

// Bottom-up for the vertices with more than MS_PULL_HUB neighbours: a work-group
// per hub, its W-lane groups striding over the list (eight neighbours per
// round as in k_ms_pull), partial words ORed in LDS.
template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_pull_hubs(MsArgs a, const uint32_t *hubs) {
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t NG = 256 / W;  // groups per work-group
  constexpr int PER = 8 / W;
  __shared__ uint32_t s_acc[W];
  __shared__ uint32_t s_seen[W];
  const uint32_t t = threadIdx.x, j = t % W, grp = t / W, lane = t & 63;
  const uint32_t u = hubs[blockIdx.x];
  if (t < W) {
    s_acc[t] = 0;
    s_seen[t] = a.seen[(size_t)u * W + t];
  }
  __syncthreads();
  const uint32_t s = s_seen[j];
  bool all = true;
#pragma unroll
  for (int w = 0; w < W; w++) all = all && s_seen[w] == ~0u;
  if (!all) {  // uniform over the work-group
    const unsigned long long gmask = ((W == 64) ? ~0ull : ((1ull << W) - 1)) << ((lane / W) * W);
    const PtrT b = col_ptr[u], e = col_ptr[u + 1];
    uint32_t acc = 0;
    for (PtrT k = b + (PtrT)grp * 8; k < e; k += (PtrT)NG * 8) {
      uint32_t ids[PER];
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const PtrT kk = k + (PtrT)(q * W + j);
        const uint32_t r = a.row_idx[kk < e ? kk : b];  // b < e: no branch around the load
        ids[q] = kk < e ? r : 0xFFFFFFFFu;
      }
      acc |= ms_gather8<W>(a, ids, (lane / W) * W, j);
      const bool full = (acc | s) == ~0u;
      if ((__ballot(!full) & gmask) == 0) break;
    }
    if (acc & ~s) atomicOr(&s_acc[j], acc & ~s);
  }
  __syncthreads();
  if (t < W) {
    const uint32_t nb = s_acc[t];
    const size_t o = (size_t)u * W + t;
    if (nb) {
      a.seen[o] = s_seen[t] | nb;
      if (a.level > 254) atomicOr(a.overflow, 1u);
      else ms_write_levels(a.lv, 32 * W, u, t, nb, (uint8_t)a.level);
    }
    a.nxt[o] = nb;
  }
  if (t == 0) {
    bool any = false;
    for (int w = 0; w < W; w++) any = any || s_acc[w] != 0;
    if (any) {
      a.q_out[atomicAdd(a.q_count_out, 1u)] = u;
      atomicAdd(a.q_edges_out, (unsigned long long)(col_ptr[u + 1] - col_ptr[u]));
      atomicOr(&a.act_next[u >> 5], 1u << (u & 31));
    }
  }
}

// vertices with more than MS_PULL_HUB neighbours (once per batch, outside the timer)
template <typename PtrT>
__global__ void k_ms_hub_list(const PtrT *col_ptr, uint32_t n_rows, uint32_t *hubs,
                              uint32_t *n_hubs) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_rows;
       u += gridDim.x * blockDim.x)
    if ((uint64_t)(col_ptr[u + 1] - col_ptr[u]) > MS_PULL_HUB) hubs[atomicAdd(n_hubs, 1u)] = u;
}

// ---- set-up ------------------------------------------------------------------
// seen/cur: source bits; bits >= n_bits pre-set in seen so that they count as
// reached; levels 0xFF everywhere, 0 at sources.
template <int W>
__global__ void k_ms_init_words(uint32_t *seen, uint32_t *cur, uint32_t *nxt, uint32_t n_cols,
                                uint32_t n_bits) {
  const size_t n = (size_t)n_cols * W;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t j = (uint32_t)(i % W);
    const uint32_t lo = 32 * j;
    uint32_t pad = 0;
    if (n_bits <= lo) pad = ~0u;
    else if (n_bits < lo + 32) pad = ~0u << (n_bits - lo);
    seen[i] = pad;
    cur[i] = 0;
    nxt[i] = 0;
  }
}

template <int W>
__global__ void k_ms_sources(const uint32_t *src, const uint16_t *bit, uint32_t n_src,
                             uint32_t *seen, uint32_t *cur, uint8_t *lv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_src) return;
  const uint32_t u = src[i], b = bit[i];
  atomicOr(&seen[(size_t)u * W + b / 32], 1u << (b % 32));
  atomicOr(&cur[(size_t)u * W + b / 32], 1u << (b % 32));
  lv[(size_t)u * (32 * W) + b] = 0;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- output: vertex-major [V][NB] -> one row per search [n_bits][V] -----------
__global__ __launch_bounds__(256) void k_ms_transpose(const uint8_t *lv, uint32_t NB,
                                                      uint32_t n_bits, uint32_t n_cols,
                                                      uint8_t *out) {
  __shared__ uint8_t tile[64][257];
  const uint32_t v0 = blockIdx.x * 64;
  // load 64 vertices x NB bytes
  for (uint32_t i = threadIdx.x; i < 64 * NB; i += 256) {
    const uint32_t dv = i / NB, b = i % NB;
    tile[dv][b] = (v0 + dv < n_cols) ? lv[(size_t)(v0 + dv) * NB + b] : 0xFF;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 64 * n_bits; i += 256) {
    const uint32_t b = i / 64, dv = i % 64;
    if (v0 + dv < n_cols) out[(size_t)b * n_cols + v0 + dv] = tile[dv][b];
  }
}

// sum over searches of the reference's edgesCount: deg(u) x searches reaching u
template <typename PtrT, int W>
__global__ void k_ms_edges(const PtrT *col_ptr, const uint32_t *seen, uint32_t n_rows,
                           uint32_t n_bits, unsigned long long *out) {
  unsigned long long acc = 0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_rows;
       u += gridDim.x * blockDim.x) {
    uint32_t c = 0;
    for (int j = 0; j < W; j++) {
      uint32_t s = seen[(size_t)u * W + j];
      const uint32_t lo = 32 * j;
      if (n_bits <= lo) s = 0;
      else if (n_bits < lo + 32) s &= ~(~0u << (n_bits - lo));
      c += __popc(s);
    }
    if (c) acc += (unsigned long long)c * (unsigned long long)(col_ptr[u + 1] - col_ptr[u]);
  }
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

// ---- host side ------------------------------------------------------------------
static unsigned ms_grid(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return g > cap ? cap : (unsigned)g;
}

template <typename PtrT, int W>
static MsResult ms_run_t(bfsx_ctx *ctx, const uint32_t *d_src, const uint16_t *d_bit,
                         uint32_t n_src, const std::vector<uint32_t> &distinct_src,
                         uint32_t n_bits, uint8_t *levels_out, uint32_t push_divisor) {
  hipStream_t s = ctx->stream;
  const uint64_t n = ctx->n_cols;
  const uint32_t NB = 32 * W;
  ensure(ctx->ms_seen, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_cur, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_nxt, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_lv, (size_t)n * NB, s);
  ensure(ctx->ms_hubs, sizeof(uint32_t) * (n + 1), s);
  ensure(ctx->ms_act, sizeof(uint32_t) * 2 * (n / 32 + 2), s);
  ensure(ctx->scratch, 64, s);
  const bool symmetric = (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols;
  uint32_t *seen = ctx->ms_seen.as<uint32_t>();
  uint32_t *cur = ctx->ms_cur.as<uint32_t>();
  uint32_t *nxt = ctx->ms_nxt.as<uint32_t>();
  uint8_t *lv = ctx->ms_lv.as<uint8_t>();
  LevelCtl *ctl = ctx->ctl.as<LevelCtl>();
  const unsigned cap = (unsigned)ctx->cu_count * 16;
  // initial state (outside the timer)
  hipLaunchKernelGGL(k_ms_init_words<W>, dim3(ms_grid(n * W, 256, cap)), dim3(256), 0, s, seen,
                     cur, nxt, (uint32_t)n, n_bits);
  BFSX_HIP(hipMemsetAsync(lv, 0xFF, (size_t)n * NB, s));
  hipLaunchKernelGGL(k_ms_sources<W>, dim3((n_src + 255) / 256), dim3(256), 0, s, d_src, d_bit,
                     n_src, seen, cur, lv);
  BFSX_HIP(hipMemcpyAsync(ctx->queue[0].p, distinct_src.data(),
                          sizeof(uint32_t) * distinct_src.size(), hipMemcpyHostToDevice, s));
  BFSX_HIP(hipMemsetAsync(ctl, 0, CTL_RESET_BYTES, s));  // (keeps the level-end counter)
  // vertices that bottom-up levels hand to k_ms_pull_hubs
  uint32_t n_phubs = 0;
  if (symmetric) {
    ensure(ctx->ms_phubs, sizeof(uint32_t) * ((uint64_t)ctx->n_rows + 1), s);
    uint32_t *d_np = ctx->scratch.as<uint32_t>() + 12;
    BFSX_HIP(hipMemsetAsync(d_np, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_ms_hub_list<PtrT>, dim3(ms_grid(ctx->n_rows, 256, cap)), dim3(256), 0, s,
                       static_cast<const PtrT *>(ctx->col_ptr.p), ctx->n_rows,
                       ctx->ms_phubs.as<uint32_t>(), d_np);
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(&n_phubs, d_np, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  BFSX_HIP(hipStreamSynchronize(s));

  MsResult r;
  BFSX_HIP(hipEventRecord(ctx->ev0, s));
  uint32_t F = (uint32_t)distinct_src.size(), qi = 0, level = 0;
  const double avg_deg = ctx->n_rows ? (double)ctx->nnz / ctx->n_rows : 0.0;
  double f_edges = F * avg_deg;  // first level: estimate; then measured
  uint32_t *hubs = ctx->ms_hubs.as<uint32_t>();
  uint32_t *n_hubs = ctx->scratch.as<uint32_t>() + 8;
  const uint32_t n_groups = (uint32_t)((n + (64 / W) - 1) / (64 / W));
  const size_t act_words = n / 32 + 2;
  uint32_t *act[2] = {ctx->ms_act.as<uint32_t>(), ctx->ms_act.as<uint32_t>() + act_words};
  int ai = 0;  // act[ai] receives the next level's active vertices
  while (F > 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(ms_grid(act_words, 256, cap)), dim3(256), 0, s, act[ai],
                       act_words, 0u);
    MsArgs a{};
    a.act_cur = level == 0 ? nullptr : act[ai ^ 1];
    a.act_next = act[ai];
    a.col_ptr = ctx->col_ptr.p;
    a.row_idx = ctx->row_idx.as<uint32_t>();
    a.n_rows = ctx->n_rows;
    a.n_cols = (uint32_t)n;
    a.seen = seen;
    a.cur = cur;
    a.nxt = nxt;
    a.lv = lv;
    a.q_in = ctx->queue[qi].as<uint32_t>();
    a.q_out = ctx->queue[qi ^ 1].as<uint32_t>();
    a.q_count_out = &ctl->q_count[qi ^ 1];
    a.F = F;
    a.level = level + 1;
    a.overflow = &ctl->overflow;
    a.q_edges_out = &ctl->q_edges[qi ^ 1];
    const bool push = !symmetric || f_edges * push_divisor < (double)ctx->nnz;
    if (!push && ctx->row_idx_pull.p) a.row_idx = ctx->row_idx_pull.as<uint32_t>();
    if (push) {
      hipLaunchKernelGGL(k_fill_u32, dim3(ms_grid(n * W, 256, cap)), dim3(256), 0, s, nxt,
                         (size_t)(n * W), 0u);
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, n_hubs, (size_t)1, 0u);
      hipLaunchKernelGGL((k_ms_push<PtrT, W>), dim3(ms_grid((uint64_t)F * 64, 256, cap)),
                         dim3(256), 0, s, a, hubs, n_hubs);
      hipLaunchKernelGGL((k_ms_push_hubs<PtrT, W>), dim3(cap), dim3(256), 0, s, a,
                         (const uint32_t *)hubs, (const uint32_t *)n_hubs);
      hipLaunchKernelGGL((k_ms_finalize<PtrT, W>), dim3(ms_grid((uint64_t)n_groups * 64, 256, cap)),
                         dim3(256), 0, s, a, n_groups);
      r.push_levels++;
    } else {
      hipLaunchKernelGGL((k_ms_pull<PtrT, W>), dim3(ms_grid((uint64_t)n_groups * 64, 256, cap)),
                         dim3(256), 0, s, a, n_groups);
      if (n_phubs)
        hipLaunchKernelGGL((k_ms_pull_hubs<PtrT, W>), dim3(n_phubs), dim3(256), 0, s, a,
                           (const uint32_t *)ctx->ms_phubs.as<uint32_t>());
      r.pull_levels++;
    }
    BFSX_HIP(hipGetLastError());
    // next frontier = nxt; reset the consumed queue counter and degree sum
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, &ctl->q_count[qi], (size_t)1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<uint32_t *>(&ctl->q_edges[qi]), (size_t)2, 0u);
    BFSX_HIP(hipMemcpyAsync(ctx->h_ctl, ctl, sizeof(LevelCtl), hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    if (ctx->h_ctl->overflow) fail(BFSX_ERR_OVERFLOW, "a level above 254 does not fit uint8 levels");
    F = ctx->h_ctl->q_count[qi ^ 1];
    f_edges = (double)ctx->h_ctl->q_edges[qi ^ 1];
    std::swap(cur, nxt);
    qi ^= 1;
    ai ^= 1;
    level++;
    r.n_levels++;
  }
  BFSX_HIP(hipEventRecord(ctx->ev1, s));
  BFSX_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  BFSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  r.ms = ms;
  // edges summed over searches, then the per-search level rows
  unsigned long long *d_e = ctx->scratch.as<unsigned long long>();
  BFSX_HIP(hipMemsetAsync(d_e, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL((k_ms_edges<PtrT, W>), dim3(ms_grid(ctx->n_rows, 256, cap)), dim3(256), 0, s,
                     static_cast<const PtrT *>(ctx->col_ptr.p), seen, ctx->n_rows, n_bits, d_e);
  unsigned long long edges = 0;
  BFSX_HIP(hipMemcpyAsync(&edges, d_e, sizeof(edges), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  r.edges = edges;
  if (levels_out) {
    DevBuf out;
    ensure(out, (size_t)n_bits * n, s);
    hipLaunchKernelGGL(k_ms_transpose, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, lv, NB,
                       n_bits, (uint32_t)n, out.as<uint8_t>());
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(levels_out, out.p, (size_t)n_bits * n, hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    release(out);
  }
  return r;
}

MsResult ms_run(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits, uint32_t n_src,
                uint32_t n_bits, uint8_t *levels_out, uint32_t push_divisor) {
  if (!ctx->has_graph) fail(BFSX_ERR_STATE, "no graph uploaded");
  if (n_bits == 0 || n_bits > 256) fail(BFSX_ERR_ARG, "n_bits must be in [1, 256]");
  if (n_src == 0) fail(BFSX_ERR_ARG, "no sources");
  std::vector<uint32_t> distinct;
  distinct.reserve(n_src);
  for (uint32_t i = 0; i < n_src; i++) {
    if (sources[i] >= ctx->n_cols) fail(BFSX_ERR_ARG, "source out of range");
    if (bits[i] >= n_bits) fail(BFSX_ERR_ARG, "source bit out of range");
    distinct.push_back(sources[i]);
  }
  std::sort(distinct.begin(), distinct.end());
  distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
  hipStream_t s = ctx->stream;
  DevBuf ds, db;
  ensure(ds, sizeof(uint32_t) * n_src, s);
  ensure(db, sizeof(uint16_t) * n_src, s);
  BFSX_HIP(hipMemcpyAsync(ds.p, sources, sizeof(uint32_t) * n_src, hipMemcpyHostToDevice, s));
  BFSX_HIP(hipMemcpyAsync(db.p, bits, sizeof(uint16_t) * n_src, hipMemcpyHostToDevice, s));
  const int W = n_bits <= 32 ? 1 : n_bits <= 64 ? 2 : n_bits <= 128 ? 4 : 8;
  MsResult r;
  try {
    const bool wide = ctx->ptr_bytes == 8;
#define BFSX_MS_CASE(WV)                                                                    \
  case WV:                                                                                  \
    r = wide ? ms_run_t<uint64_t, WV>(ctx, ds.as<uint32_t>(), db.as<uint16_t>(), n_src,     \
                                      distinct, n_bits, levels_out, push_divisor)           \
             : ms_run_t<uint32_t, WV>(ctx, ds.as<uint32_t>(), db.as<uint16_t>(), n_src,     \
                                      distinct, n_bits, levels_out, push_divisor);          \
    break;
    switch (W) {
      BFSX_MS_CASE(1)
      BFSX_MS_CASE(2)
      BFSX_MS_CASE(4)
      BFSX_MS_CASE(8)
    }
#undef BFSX_MS_CASE
  } catch (...) {
    release(ds);
    release(db);
    throw;
  }
  release(ds);
  release(db);
  return r;
}

}  // namespace bfsx
