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
//
// The level loop is device-driven: K levels are captured once into a
// hipGraph and replayed.  Each level's kernels read the frontier size, the
// level and the direction from the control block; a one-wave kernel at the
// start of a level picks the direction (top-down while the frontier's edges
// times ms_push_divisor stay below nnz), one at the end counts, sets the
// termination flag and writes the level-end words the host polls once per
// launch.  After the last level every kernel of the remaining captured
// levels returns at once.
//
// On symmetric graphs the searches run on a copy with vertices renamed by
// degree (hubs first, made once per graph outside the timed region): a
// bottom-up level gathers the 32-byte bitsets of neighbours, most of which
// are hubs, and those then share a few L2-resident megabytes.  Sources are
// renamed on the way in and levels on the way out.
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
  const uint32_t *row_idx;       // lists of top-down levels
  const uint32_t *row_idx_pull;  // lists of bottom-up levels (hub first)
  uint32_t n_rows, n_cols;
  uint32_t *seen, *cur, *nxt;
  uint8_t *lv;
  uint32_t *q_in, *q_out, *q_count_out;
  uint32_t qi;     // queue buffer holding this level's active vertices
  uint32_t F;      // (set from the control block at kernel start)
  uint32_t level;  // level being written (the new bits' level), likewise
  uint32_t *overflow;
  unsigned long long *q_edges_out;  // out-degree sum of the next active queue
  uint32_t pull_hub;  // bottom-up: longer lists go to k_ms_pull_hubs
  LevelCtl *ctl;
  // one bit per vertex: in some search's frontier (cur != 0); not used on
  // the first level.  Bottom-up levels probe it (a few MiB, L2 resident)
  // before reading a neighbour's bitset, whose 32 bytes cost a whole line.
  const uint32_t *act_cur;
  uint32_t *act_next;
};

// Reads this level's state from the control block into a; false when the
// search has ended or the level runs in the other direction (uniform over
// the launch, so every thread returns together).
__device__ __forceinline__ bool ms_level(MsArgs &a, uint32_t pull) {
  const LevelCtl *c = a.ctl;
  if (c->done || c->choice != pull) return false;
  a.F = c->q_count[a.qi];
  a.level = c->level + 1;
  if (c->level == 0) a.act_cur = nullptr;
  return true;
}

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
// Bottom-up levels hand vertices with more neighbours than a.pull_hub to
// k_ms_pull_hubs (a work-group each): a W-lane group scanning a long list
// alone would hold its wave for hundreds of rounds.

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
__global__ __launch_bounds__(256) void k_ms_push(MsArgs a, uint32_t *hubs) {
  if (!ms_level(a, 0)) return;
  uint32_t *n_hubs = &a.ctl->ms_hub_count;
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
__global__ __launch_bounds__(256) void k_ms_push_hubs(MsArgs a, const uint32_t *hubs) {
  if (!ms_level(a, 0)) return;
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  constexpr uint32_t G = 64 / W;
  const uint32_t grp = (threadIdx.x & 63) / W;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * 4;
  const uint32_t nh = a.ctl->ms_hub_count;
  for (uint32_t h = 0; h < nh; h++)
    ms_push_edges<PtrT, W>(a, col_ptr, hubs[h], (PtrT)(wave * G + grp), (PtrT)nwaves * G);
}

// ---- after a top-down level: keep unseen bits, levels, next active queue ----
// Also clears this level's frontier words (cur), which become the next
// level's nxt: a following top-down level finds them zero.
template <typename PtrT, int W>
__global__ __launch_bounds__(256) void k_ms_finalize(MsArgs a, uint32_t n_groups_total) {
  if (!ms_level(a, 0)) return;
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
      if (a.cur[o]) a.cur[o] = 0u;
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
  if (!ms_level(a, 1)) return;
  a.row_idx = a.row_idx_pull;
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
    const bool hub = (uint64_t)(e - b) > a.pull_hub;
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
  if (!ms_level(a, 1)) return;
  a.row_idx = a.row_idx_pull;
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

// vertices with more than pull_hub neighbours (once per batch, outside the timer)
template <typename PtrT>
__global__ void k_ms_hub_list(const PtrT *col_ptr, uint32_t n_rows, uint32_t pull_hub,
                              uint32_t *hubs, uint32_t *n_hubs) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_rows;
       u += gridDim.x * blockDim.x)
    if ((uint64_t)(col_ptr[u + 1] - col_ptr[u]) > pull_hub) hubs[atomicAdd(n_hubs, 1u)] = u;
}

// ---- level start and end (one wave each) -------------------------------------
// Direction of the level, reset of the output queue's counters.
__global__ void k_ms_begin(LevelCtl *ctl, uint32_t qi, uint64_t nnz, uint32_t divisor,
                           int symmetric) {
  if (threadIdx.x != 0 || ctl->done) return;
  const uint32_t qo = qi ^ 1;
  const bool push = !symmetric || ctl->q_edges[qi] * (unsigned long long)divisor < nnz;
  ctl->ms_prev_pull = ctl->choice;
  ctl->choice = push ? 0u : 1u;
  ctl->ms_hub_count = 0;
  ctl->q_count[qo] = 0;
  ctl->q_edges[qo] = 0;
}

// Clears the next-active map, and the nxt words when a top-down level
// follows a bottom-up one (a bottom-up level writes every nxt word; a
// top-down level's finalize leaves the next top-down level's nxt zero).
__global__ void k_ms_clear(const LevelCtl *ctl, uint32_t *act_next, size_t act_words,
                           uint32_t *nxt, size_t nxt_words) {
  if (ctl->done) return;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (size_t i = i0; i < act_words; i += stride) act_next[i] = 0u;
  if (ctl->choice == 0 && ctl->ms_prev_pull)
    for (size_t i = i0; i < nxt_words; i += stride) nxt[i] = 0u;
}

// Counts, termination, and the tagged level-end words for the host.
__global__ void k_ms_end(LevelCtl *ctl, uint32_t qi, LevelMirror *mirror) {
  if (threadIdx.x != 0) return;
  const uint32_t qo = qi ^ 1;
  if (!ctl->done) {
    ctl->kind_count[ctl->choice ? 1 : 0]++;
    ctl->level += 1;
    ctl->done = ctl->q_count[qo] == 0;
  }
  const uint32_t seq = ctl->seq + 1;
  ctl->seq = seq;
  const unsigned long long t32 = (unsigned long long)seq << 32;
  const unsigned long long t16 = (unsigned long long)(seq & 0xFFFFu) << 48;
  mirror += seq & 1u;  // the set for this parity
  mirror->w[0] = t32 | ctl->q_count[qo];
  mirror->w[1] = t16 | (ctl->q_edges[qo] & MIRROR_48);
  mirror->w[2] = t16 | (ctl->overflow ? 1ull : 0ull);
  mirror->w[3] = t32 | (ctl->choice & 0xFFu) | ((unsigned long long)(ctl->done ? 1u : 0u) << 8);
  __threadfence_system();
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

// ---- renaming sources and levels ---------------------------------------------
__global__ void k_ms_map(uint32_t *v, uint32_t n, const uint32_t *rank) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = rank[v[i]];
}

// lv_out[v] = lv[rank[v]] (NB bytes per vertex, as 16-byte words)
__global__ void k_ms_unrename(const uint4 *lv, uint32_t words_per_vertex, uint32_t n,
                              const uint32_t *rank, uint4 *lv_out) {
  const uint64_t total = (uint64_t)n * words_per_vertex;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = i / words_per_vertex, k = i % words_per_vertex;
    lv_out[i] = lv[(uint64_t)rank[v] * words_per_vertex + k];
  }
}

// sum of out-degree of the distinct sources (the first level's frontier edges)
template <typename PtrT>
__global__ void k_ms_src_edges(const PtrT *col_ptr, uint32_t n_rows, const uint32_t *q,
                               uint32_t F, unsigned long long *out) {
  unsigned long long d = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < F; i += gridDim.x * blockDim.x)
    if (q[i] < n_rows) d += (unsigned long long)(col_ptr[q[i] + 1] - col_ptr[q[i]]);
  for (int k = 32; k > 0; k >>= 1) d += __shfl_xor(d, k, 64);
  if ((threadIdx.x & 63) == 0 && d) atomicAdd(out, d);
}

// ---- host side ------------------------------------------------------------------
static unsigned ms_grid(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return g > cap ? cap : (unsigned)g;
}

static void ms_drop_loop(bfsx_ctx *ctx) {
  if (ctx->ms_exec) (void)hipGraphExecDestroy(ctx->ms_exec);
  if (ctx->ms_graph) (void)hipGraphDestroy(ctx->ms_graph);
  ctx->ms_exec = nullptr;
  ctx->ms_graph = nullptr;
  ctx->ms_key = 0;
}

void ms_drop(bfsx_ctx *ctx) {
  ms_drop_loop(ctx);
  for (DevBuf *b : {&ctx->ms_rcp, &ctx->ms_rri, &ctx->ms_rank, &ctx->ms_order}) release(*b);
  ctx->ms_relabelled = false;
}

// The graph the batch searches: the relabelled copy (made on first use) or
// the resident one.
struct MsGraph {
  const void *col_ptr;
  const uint32_t *row_idx, *row_idx_pull;
  const uint32_t *rank;  // old id -> searched id, or null
  const uint32_t *order; // searched id -> old id, or null
};

static MsGraph ms_graph_of(bfsx_ctx *ctx, bool symmetric) {
  MsGraph g{ctx->col_ptr.p, ctx->row_idx.as<uint32_t>(), ctx->row_idx.as<uint32_t>(), nullptr,
            nullptr};
  if (ctx->row_idx_pull.p) g.row_idx_pull = ctx->row_idx_pull.as<uint32_t>();
  if (!symmetric || !ctx->params.ms_relabel) return g;
  if (!ctx->ms_relabelled) {
    hipStream_t s = ctx->stream;
    const uint64_t n = ctx->n_cols;
    ensure(ctx->ms_rank, sizeof(uint32_t) * std::max<uint64_t>(n, 1), s);
    ensure(ctx->ms_order, sizeof(uint32_t) * std::max<uint64_t>(n, 1), s);
    ensure(ctx->ms_rcp, (size_t)ctx->ptr_bytes * (n + 1), s);
    ensure(ctx->ms_rri, sizeof(uint32_t) * std::max<uint64_t>(ctx->nnz, 1), s);
    relabel_by_degree(ctx, ctx->ms_rank.as<uint32_t>(), ctx->ms_order.as<uint32_t>(),
                      ctx->ms_rcp.p, ctx->ms_rri.as<uint32_t>());
    ctx->ms_relabelled = true;
  }
  // relabelled lists ascend in new ids, which puts hubs first: one copy
  // serves both directions
  g.col_ptr = ctx->ms_rcp.p;
  g.row_idx = g.row_idx_pull = ctx->ms_rri.as<uint32_t>();
  g.rank = ctx->ms_rank.as<uint32_t>();
  g.order = ctx->ms_order.as<uint32_t>();
  return g;
}

template <typename PtrT, int W>
static MsResult ms_run_t(bfsx_ctx *ctx, const uint32_t *d_src, const uint16_t *d_bit,
                         uint32_t n_src, const std::vector<uint32_t> &distinct_src,
                         uint32_t n_bits, uint8_t *levels_out) {
  hipStream_t s = ctx->stream;
  const uint64_t n = ctx->n_cols;
  const uint32_t NB = 32 * W;
  const bool symmetric = (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols;
  const MsGraph G = ms_graph_of(ctx, symmetric);
  const PtrT *gcp = static_cast<const PtrT *>(G.col_ptr);
  ensure(ctx->ms_seen, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_cur, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_nxt, sizeof(uint32_t) * n * W, s);
  ensure(ctx->ms_lv, (size_t)n * NB, s);
  ensure(ctx->ms_hubs, sizeof(uint32_t) * (n + 1), s);
  const size_t act_words = n / 32 + 2;
  ensure(ctx->ms_act, sizeof(uint32_t) * 2 * act_words, s);
  ensure(ctx->scratch, 64, s);
  uint32_t *seen = ctx->ms_seen.as<uint32_t>();
  uint32_t *buf[2] = {ctx->ms_cur.as<uint32_t>(), ctx->ms_nxt.as<uint32_t>()};
  uint32_t *act[2] = {ctx->ms_act.as<uint32_t>(), ctx->ms_act.as<uint32_t>() + act_words};
  uint8_t *lv = ctx->ms_lv.as<uint8_t>();
  LevelCtl *ctl = ctx->ctl.as<LevelCtl>();
  const unsigned cap = (unsigned)ctx->cu_count * 16;
  const uint32_t pull_hub = ctx->params.ms_pull_hub;

  // initial state (outside the timer): sources renamed, bitsets, levels,
  // the first queue and its edge count, the control block (its level-end
  // counter kept)
  uint32_t *src = const_cast<uint32_t *>(d_src);
  uint32_t *q0 = ctx->queue[0].as<uint32_t>();
  const uint32_t F0 = (uint32_t)distinct_src.size();
  BFSX_HIP(hipMemcpyAsync(q0, distinct_src.data(), sizeof(uint32_t) * F0, hipMemcpyHostToDevice, s));
  if (G.rank) {
    hipLaunchKernelGGL(k_ms_map, dim3((n_src + 255) / 256), dim3(256), 0, s, src, n_src, G.rank);
    hipLaunchKernelGGL(k_ms_map, dim3((F0 + 255) / 256), dim3(256), 0, s, q0, F0, G.rank);
  }
  hipLaunchKernelGGL(k_ms_init_words<W>, dim3(ms_grid(n * W, 256, cap)), dim3(256), 0, s, seen,
                     buf[0], buf[1], (uint32_t)n, n_bits);
  BFSX_HIP(hipMemsetAsync(lv, 0xFF, (size_t)n * NB, s));
  hipLaunchKernelGGL(k_ms_sources<W>, dim3((n_src + 255) / 256), dim3(256), 0, s, src, d_bit,
                     n_src, seen, buf[0], lv);
  hipLaunchKernelGGL(k_fill_u32, dim3(ms_grid(2 * act_words, 256, cap)), dim3(256), 0, s,
                     act[0], 2 * act_words, 0u);
  BFSX_HIP(hipMemsetAsync(ctl, 0, CTL_RESET_BYTES, s));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, &ctl->q_count[0], (size_t)1, F0);
  hipLaunchKernelGGL((k_ms_src_edges<PtrT>), dim3(ms_grid(F0, 256, 64)), dim3(256), 0, s, gcp,
                     ctx->n_rows, q0, F0, &ctl->q_edges[0]);
  // vertices that bottom-up levels hand to k_ms_pull_hubs
  uint32_t n_phubs = 0;
  if (symmetric) {
    ensure(ctx->ms_phubs, sizeof(uint32_t) * ((uint64_t)ctx->n_rows + 1), s);
    uint32_t *d_np = ctx->scratch.as<uint32_t>() + 12;
    BFSX_HIP(hipMemsetAsync(d_np, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_ms_hub_list<PtrT>, dim3(ms_grid(ctx->n_rows, 256, cap)), dim3(256), 0, s,
                       gcp, ctx->n_rows, pull_hub, ctx->ms_phubs.as<uint32_t>(), d_np);
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(&n_phubs, d_np, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));

  // K device-driven levels, captured once per shape
  const uint32_t K = ctx->params.ms_graph_levels;
  const uint32_t n_groups = (uint32_t)((n + (64 / W) - 1) / (64 / W));
  uint64_t key = 1469598103934665603ull;
  for (uint64_t v : {(uint64_t)W, (uint64_t)sizeof(PtrT), (uint64_t)K, (uint64_t)n,
                     (uint64_t)n_phubs, (uint64_t)pull_hub, (uint64_t)ctx->params.ms_push_divisor,
                     (uint64_t)symmetric, (uint64_t)(uintptr_t)G.col_ptr,
                     (uint64_t)(uintptr_t)G.row_idx_pull, (uint64_t)(uintptr_t)seen,
                     (uint64_t)(uintptr_t)buf[0], (uint64_t)(uintptr_t)buf[1],
                     (uint64_t)(uintptr_t)lv, (uint64_t)(uintptr_t)act[0],
                     (uint64_t)(uintptr_t)ctx->ms_hubs.p, (uint64_t)(uintptr_t)ctx->ms_phubs.p,
                     (uint64_t)(uintptr_t)ctx->queue[0].p, (uint64_t)(uintptr_t)ctx->queue[1].p})
    key = (key ^ v) * 1099511628211ull;
  if (!ctx->ms_exec || ctx->ms_key != key) {
    ms_drop_loop(ctx);
    BFSX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      for (uint32_t k = 0; k < K; k++) {
        const uint32_t p = k & 1u;
        MsArgs a{};
        a.col_ptr = G.col_ptr;
        a.row_idx = G.row_idx;
        a.row_idx_pull = G.row_idx_pull;
        a.n_rows = ctx->n_rows;
        a.n_cols = (uint32_t)n;
        a.seen = seen;
        a.cur = buf[p];
        a.nxt = buf[p ^ 1];
        a.lv = lv;
        a.qi = p;
        a.q_in = ctx->queue[p].as<uint32_t>();
        a.q_out = ctx->queue[p ^ 1].as<uint32_t>();
        a.q_count_out = &ctl->q_count[p ^ 1];
        a.q_edges_out = &ctl->q_edges[p ^ 1];
        a.overflow = &ctl->overflow;
        a.pull_hub = pull_hub;
        a.ctl = ctl;
        a.act_cur = act[p ^ 1];
        a.act_next = act[p];
        hipLaunchKernelGGL(k_ms_begin, dim3(1), dim3(64), 0, s, ctl, p, (uint64_t)ctx->nnz,
                           ctx->params.ms_push_divisor, symmetric ? 1 : 0);
        hipLaunchKernelGGL(k_ms_clear, dim3(cap), dim3(256), 0, s, (const LevelCtl *)ctl, act[p],
                           act_words, buf[p ^ 1], (size_t)(n * W));
        hipLaunchKernelGGL((k_ms_push<PtrT, W>), dim3(cap), dim3(256), 0, s, a,
                           ctx->ms_hubs.as<uint32_t>());
        hipLaunchKernelGGL((k_ms_push_hubs<PtrT, W>), dim3(cap), dim3(256), 0, s, a,
                           (const uint32_t *)ctx->ms_hubs.as<uint32_t>());
        hipLaunchKernelGGL((k_ms_finalize<PtrT, W>), dim3(ms_grid((uint64_t)n_groups * 64, 256, cap)),
                           dim3(256), 0, s, a, n_groups);
        if (symmetric) {
          hipLaunchKernelGGL((k_ms_pull<PtrT, W>), dim3(ms_grid((uint64_t)n_groups * 64, 256, cap)),
                             dim3(256), 0, s, a, n_groups);
          if (n_phubs)
            hipLaunchKernelGGL((k_ms_pull_hubs<PtrT, W>), dim3(n_phubs), dim3(256), 0, s, a,
                               (const uint32_t *)ctx->ms_phubs.as<uint32_t>());
        }
        hipLaunchKernelGGL(k_ms_end, dim3(1), dim3(64), 0, s, ctl, p, ctx->d_mirror);
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    BFSX_HIP(hipStreamEndCapture(s, &ctx->ms_graph));
    BFSX_HIP(hipGraphInstantiate(&ctx->ms_exec, ctx->ms_graph, nullptr, nullptr, 0));
    ctx->ms_key = key;
  }

  // the timed level loop: one graph launch per K levels, the termination
  // flag read from the tagged level-end words
  MsResult r;
  BFSX_HIP(hipMemcpyAsync(ctx->h_ctl, ctl, sizeof(LevelCtl), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  uint32_t seq = ctx->h_ctl->seq;
  BFSX_HIP(hipEventRecord(ctx->ev0, s));
  const uint64_t max_launches = n / K + 4;  // a search has at most n levels
  for (uint64_t launches = 0;; launches++) {
    if (launches > max_launches) fail(BFSX_ERR_DEVICE, "batched level loop did not terminate");
    BFSX_HIP(hipGraphLaunch(ctx->ms_exec, s));
    seq += K;
    read_level_end(ctx, seq, 0);
    if (ctx->view.done) break;
  }
  BFSX_HIP(hipEventRecord(ctx->ev1, s));
  BFSX_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  BFSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  r.ms = ms;
  BFSX_HIP(hipMemcpyAsync(ctx->h_ctl, ctl, sizeof(LevelCtl), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  if (ctx->h_ctl->overflow) fail(BFSX_ERR_OVERFLOW, "a level above 254 does not fit uint8 levels");
  r.push_levels = ctx->h_ctl->kind_count[0];
  r.pull_levels = ctx->h_ctl->kind_count[1];
  r.n_levels = r.push_levels + r.pull_levels;

  // edges summed over searches, then the per-search level rows
  unsigned long long *d_e = ctx->scratch.as<unsigned long long>();
  BFSX_HIP(hipMemsetAsync(d_e, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL((k_ms_edges<PtrT, W>), dim3(ms_grid(ctx->n_rows, 256, cap)), dim3(256), 0, s,
                     gcp, seen, ctx->n_rows, n_bits, d_e);
  unsigned long long edges = 0;
  BFSX_HIP(hipMemcpyAsync(&edges, d_e, sizeof(edges), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  r.edges = edges;
  if (levels_out) {
    DevBuf out, plain;
    const uint8_t *rows = lv;
    if (G.rank) {  // back to the caller's vertex ids
      ensure(plain, (size_t)n * NB, s);
      hipLaunchKernelGGL(k_ms_unrename, dim3(ms_grid(n * (NB / 16), 256, cap)), dim3(256), 0, s,
                         reinterpret_cast<const uint4 *>(lv), NB / 16, (uint32_t)n, G.rank,
                         plain.as<uint4>());
      rows = plain.as<uint8_t>();
    }
    ensure(out, (size_t)n_bits * n, s);
    hipLaunchKernelGGL(k_ms_transpose, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, rows, NB,
                       n_bits, (uint32_t)n, out.as<uint8_t>());
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(levels_out, out.p, (size_t)n_bits * n, hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    release(out);
    release(plain);
  }
  return r;
}

MsResult ms_run(bfsx_ctx *ctx, const uint32_t *sources, const uint16_t *bits, uint32_t n_src,
                uint32_t n_bits, uint8_t *levels_out) {
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
                                      distinct, n_bits, levels_out)                         \
             : ms_run_t<uint32_t, WV>(ctx, ds.as<uint32_t>(), db.as<uint16_t>(), n_src,     \
                                      distinct, n_bits, levels_out);                        \
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
