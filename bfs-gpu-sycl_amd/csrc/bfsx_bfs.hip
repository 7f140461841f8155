// Level-synchronous BFS step kernels for gfx950 (wave64).
//
// Reference path (src/run_bfs_gpu.hpp): per level an explore kernel marks
// neighbours of the frontier in byte masks (LLB/SLB :388-477, GLB :175-385)
// and a level-generation kernel scans all V vertices to turn marks into the
// next frontier (:481-521).  Here discovery is exact-once through a visited
// bitmap with atomicOr, and the discovering lane writes the level and
// appends the vertex to the next frontier in the same kernel: one ballot and
// one atomicAdd per wave give every new vertex its slot (mbcnt prefix), so
// there is no O(V) level-generation pass.
//
// Step kinds:
//   LLB   a work-group takes WG consecutive frontier entries, scans their
//         degrees in registers/LDS and expands all their edges cooperatively
//         (binary search of the edge index over the LDS scan) -- the
//         reference's local load balancing.  Hubs above `hub_degree` are
//         taken out of the scan and expanded by a whole wave each.
//   SLB   the same kernel with the strided entry mapping of
//         run_bfs_gpu.hpp:416-417 (group + lane * n_groups).
//   GLB   device-wide scan of frontier degrees, then a persistent grid pulls
//         edge chunks (smaller ones for a frontier with few edges) from an
//         atomic work queue; each chunk finds
//         its frontier entries from a precomputed chunk->entry table and
//         stages them in LDS (no global binary search per edge).
//   PULL  bottom-up: every unvisited vertex looks for a neighbour in the
//         frontier bitmap and stops at the first hit; a wave owns 64
//         consecutive vertices, so bitmap words are written without atomics.
#include <algorithm>

#include <hip/hip_cooperative_groups.h>

#include "bfsx_internal.hpp"
#include "bfsx_mlp.hpp"

namespace bfsx {

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int l = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d, 64);
    if (l >= d) x += y;
  }
  return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Exclusive scan over a work-group; s_wave holds WG/64 + 1 entries.
template <int WG, class T>
__device__ __forceinline__ T block_excl_scan(T x, T &total, T *s_wave) {
  constexpr int NW = WG / 64;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const T inc = wave_incl_scan(x);
  if (NW == 1) {
    total = __shfl(inc, 63, 64);
    return inc - x;
  }
  if (l == 63) s_wave[w] = inc;
  __syncthreads();
  if (w == 0) {
    const T v = (l < NW) ? s_wave[l] : T(0);
    const T vi = wave_incl_scan(v);
    if (l < NW) s_wave[l] = vi - v;
    if (l == NW - 1) s_wave[NW] = vi;
  }
  __syncthreads();
  total = s_wave[NW];
  return s_wave[w] + inc - x;
}

// Exact-once discovery: a plain read filters vertices already visited (the
// bitmap is small enough to stay in L2/MALL), the atomicOr decides the winner.
__device__ __forceinline__ bool try_visit(uint32_t *vis, uint32_t n) {
  const uint32_t w = n >> 5, bit = 1u << (n & 31);
  if (vis[w] & bit) return false;
  return !(atomicOr(&vis[w], bit) & bit);
}

// Wave-aggregated append: one atomicAdd per wave, slots by mbcnt prefix.
// Every active lane of the wave must call it.
__device__ __forceinline__ void wave_append(bool isnew, uint32_t n, uint32_t *q_out,
                                            uint32_t *q_count) {
  const unsigned long long m = __ballot(isnew);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(q_count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (isnew) q_out[base + lanes_below(m)] = n;
}

// Adds a per-thread value over the work-group to *dst with one atomic.
// Every thread of the group must call it.
template <int WG>
__device__ __forceinline__ void block_atomic_add(unsigned long long v, unsigned long long *dst) {
  __shared__ unsigned long long s_part[WG / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < WG / 64; w++) t += s_part[w];
    if (t) atomicAdd(dst, t);
  }
}

// Out-degree of vertex v (v below the padded bitmap length) from the 16-bit
// array, col_ptr only when it saturated.
template <typename PtrT>
__device__ __forceinline__ unsigned long long degree16(const uint16_t *deg16, const PtrT *col_ptr,
                                                       uint32_t v) {
  const uint32_t d = deg16[v];
  return d != DEG16_ESCAPE ? (unsigned long long)d
                           : (unsigned long long)(col_ptr[v + 1] - col_ptr[v]);
}

// Out-degree of a discovered vertex (for the next frontier's edge count).
template <typename PtrT>
__device__ __forceinline__ unsigned long long out_degree(const PtrT *col_ptr, uint32_t n,
                                                         uint32_t n_rows) {
  return n < n_rows ? (unsigned long long)(col_ptr[n + 1] - col_ptr[n]) : 0ull;
}

struct KArgs {
  const void *col_ptr;
  const uint32_t *row_idx;
  uint32_t n_rows, n_cols;
  const uint32_t *q_in;
  uint32_t F;
  uint32_t *q_out;
  uint32_t *q_out_count;
  unsigned long long *q_out_edges;  // nullable
  uint32_t *vis;
  int32_t *levels;
  uint32_t *parents;  // nullable
  int32_t next_level;
  uint32_t hub_degree;
  uint8_t *mark;      // non-null: mark-then-compact mode
  unsigned long long *examined;  // nullable: bottom-up adjacency entries read
  // mark mode only: row_idx is the hot copy; a flagged entry names hot
  // vertex hot_id[h], whose visited bit is bit h of hot_vis (null: plain ids)
  const uint32_t *hot_vis;
  const uint32_t *hot_id;
};

// Visited-bitmap word and bit index of a row_idx entry x: with HOT (the hot
// copy in row_idx) a flagged x is a hot index, probed in a.hot_vis.
template <bool HOT>
__device__ __forceinline__ const uint32_t *probe_word(const KArgs &a, uint32_t x,
                                                      uint32_t &idx) {
  if (!HOT) {
    idx = x;
    return a.vis + (x >> 5);
  }
  const bool hot = (x & PULL_HUB_FLAG) != 0;
  idx = hot ? x & ~PULL_HUB_FLAG : x;
  return (hot ? a.hot_vis : a.vis) + (idx >> 5);
}

// The vertex id of a row_idx entry.
template <bool HOT>
__device__ __forceinline__ uint32_t entry_vertex(const KArgs &a, uint32_t x) {
  return (HOT && (x & PULL_HUB_FLAG)) ? a.hot_id[x & ~PULL_HUB_FLAG] : x;
}

// Two ways to discover a neighbour n of a frontier vertex:
//  * exact-once: atomicOr on the visited bitmap picks one winner, which
//    writes the level and appends n right away (few edges: thin levels);
//  * mark: a plain byte store into `mark` for every edge that sees n
//    unvisited in the bitmap as it was at the start of the level; all
//    writers store the same value, so no atomic is needed.  k_compact then
//    turns marks into levels, bitmap bits and the next queue (heavy levels,
//    where per-edge device-scope atomics would dominate).
template <typename PtrT, bool HOT>
__device__ __forceinline__ void discover(const KArgs &a, const PtrT *col_ptr, uint32_t n,
                                         uint32_t from, unsigned long long &dacc) {
  if (a.mark) {
    uint32_t idx;
    if (!((*probe_word<HOT>(a, n, idx) >> (idx & 31)) & 1u)) {
      const uint32_t v = entry_vertex<HOT>(a, n);
      a.mark[v] = 1;
      if (a.parents) a.parents[v] = from;
    }
    return;
  }
  const bool isnew = try_visit(a.vis, n);
  if (isnew) {
    a.levels[n] = a.next_level;
    if (a.parents) a.parents[n] = from;
    if (a.q_out_edges) dacc += out_degree(col_ptr, n, a.n_rows);
  }
  wave_append(isnew, n, a.q_out, a.q_out_count);
}

// Expands the edge range [lo, hi) of a work-group whose frontier entries are
// staged in LDS: s_scan[j] = first edge of slot j (exclusive scan; slots past
// the valid ones hold values >= hi), s_start[j] = its first row_idx offset.
// Each lane takes EPT consecutive edges per round: one binary search finds
// the owner slot of the first, the others advance it; all row_idx loads,
// then all visited-bitmap loads are issued before any store, so every lane
// keeps EPT independent memory chains in flight.  Exact-once discoveries
// are appended with one atomicAdd per wave and edge group, or, with
// BLOCK_APPEND on a range of at most BLOCK_APPEND_ROUNDS rounds, one per
// work-group and round (a block scan gives every thread its slots): on thin
// levels spread over many small LLB tiles, such as a grid's, per-wave
// atomics on the one queue counter serialise; on a tile holding a hub's
// list, hundreds of rounds, the block scan's barriers would make every
// round wait for the group's slowest wave, so such ranges (and GLB's)
// append per wave.  Every thread of the group calls this with the same lo
// and hi.
constexpr int EPT = 4;
constexpr int BLOCK_APPEND_ROUNDS = 2;

// Row_idx entries of the EPT consecutive edges from `base` (those below hi),
// and the frontier vertices they leave from.
template <typename PtrT, int WG>
__device__ __forceinline__ void load_edges(const KArgs &a, const PtrT *s_scan,
                                           const PtrT *s_start, const uint32_t *s_v,
                                           PtrT base, PtrT hi, uint32_t (&n)[EPT],
                                           uint32_t (&from)[EPT], bool (&valid)[EPT]) {
  uint32_t j = 0;
#pragma unroll
  for (uint32_t step = WG / 2; step > 0; step >>= 1)
    if (s_scan[j + step] <= base) j += step;
  // (the loads stay under their branches: measured faster for LLB than
  // clamped unconditional loads, whose extra reads and later issue cost
  // more than the overlap gained)
#pragma unroll
  for (int k = 0; k < EPT; k++) {
    const PtrT e = base + k;
    valid[k] = e < hi;
    while (valid[k] && j + 1 < WG && s_scan[j + 1] <= e) j++;
    n[k] = valid[k] ? a.row_idx[s_start[j] + (e - s_scan[j])] : 0u;
    from[k] = s_v[j];
  }
}

// STAGE (GLB): exact-once discoveries go to the work-group's LDS buffer
// s_buf (one LDS atomic per wave and edge group, on *s_n), which the caller
// flushes to the queue with one global atomic per chunk.
template <typename PtrT, int WG, bool BLOCK_APPEND, bool HOT, bool STAGE = false>
__device__ __forceinline__ void expand_edges(const KArgs &a, const PtrT *col_ptr,
                                             const PtrT *s_scan, const PtrT *s_start,
                                             const uint32_t *s_v, PtrT lo, PtrT hi,
                                             unsigned long long &dacc,
                                             uint32_t *s_buf = nullptr, uint32_t *s_n = nullptr) {
  constexpr PtrT PER_ROUND = (PtrT)WG * EPT;
  if (BLOCK_APPEND && !a.mark && hi > lo && hi - lo <= PER_ROUND * BLOCK_APPEND_ROUNDS) {
    // (uniform over the group: lo, hi and the mode are)
    __shared__ uint32_t s_wsum[WG / 64 + 1];
    __shared__ uint32_t s_qbase;
    for (PtrT r0 = lo; r0 < hi; r0 += PER_ROUND) {
      uint32_t n[EPT], from[EPT];
      bool valid[EPT];
      load_edges<PtrT, WG>(a, s_scan, s_start, s_v, r0 + (PtrT)threadIdx.x * EPT, hi, n, from,
                           valid);
      // The EPT discoveries of a lane as three waves of memory operations
      // instead of a chain: every visited word, then an atomicOr for each
      // target that looked unvisited, then (tracking edges) the degrees of
      // those targets, all issued before the first result is used (each
      // read would otherwise wait for the previous edge's atomic, which may
      // alias it).  Two edges of one lane to one vertex stay exact-once:
      // a thread's atomics on one address take effect in order.
      uint32_t word[EPT], bit[EPT], old[EPT];
      bool cand[EPT], isnew[EPT];
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        bit[k] = 1u << (n[k] & 31);
        word[k] = valid[k] ? a.vis[n[k] >> 5] : ~0u;
      }
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        cand[k] = valid[k] && !(word[k] & bit[k]);
        old[k] = cand[k] ? atomicOr(&a.vis[n[k] >> 5], bit[k]) : bit[k];
      }
      unsigned long long deg[EPT];
#pragma unroll
      for (int k = 0; k < EPT; k++)
        deg[k] = (a.q_out_edges && cand[k]) ? out_degree(col_ptr, n[k], a.n_rows) : 0ull;
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        isnew[k] = !(old[k] & bit[k]);
        if (isnew[k]) {
          a.levels[n[k]] = a.next_level;
          if (a.parents) a.parents[n[k]] = from[k];
          dacc += deg[k];
          cnt++;
        }
      }
      uint32_t total;
      const uint32_t ex = block_excl_scan<WG>(cnt, total, s_wsum);
      if (threadIdx.x == 0 && total) s_qbase = atomicAdd(a.q_out_count, total);
      __syncthreads();
      if (cnt) {
        uint32_t p = s_qbase + ex;
#pragma unroll
        for (int k = 0; k < EPT; k++)
          if (isnew[k]) a.q_out[p++] = n[k];
      }
      __syncthreads();  // s_qbase and s_wsum are rewritten by the next round
    }
    return;
  }
  for (PtrT base = lo + (PtrT)threadIdx.x * EPT; base < hi; base += PER_ROUND) {
    uint32_t n[EPT], from[EPT];
    bool valid[EPT];
    load_edges<PtrT, WG>(a, s_scan, s_start, s_v, base, hi, n, from, valid);
    if (a.mark) {
      uint32_t w[EPT], idx[EPT];
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        const uint32_t *p = probe_word<HOT>(a, n[k], idx[k]);
        w[k] = valid[k] ? *p : ~0u;
      }
#pragma unroll
      for (int k = 0; k < EPT; k++)
        if (valid[k] && !((w[k] >> (idx[k] & 31)) & 1u)) {
          const uint32_t v = entry_vertex<HOT>(a, n[k]);
          a.mark[v] = 1;
          if (a.parents) a.parents[v] = from[k];
        }
    } else {
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        const bool isnew = valid[k] && try_visit(a.vis, n[k]);
        if (isnew) {
          a.levels[n[k]] = a.next_level;
          if (a.parents) a.parents[n[k]] = from[k];
          if (a.q_out_edges) dacc += out_degree(col_ptr, n[k], a.n_rows);
        }
        if (STAGE) wave_append(isnew, n[k], s_buf, s_n);  // (LDS)
        else wave_append(isnew, n[k], a.q_out, a.q_out_count);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// LLB / SLB
// ---------------------------------------------------------------------------
template <typename PtrT, int WG>
struct LlbShared {
  PtrT start[WG];
  PtrT scan[WG];
  PtrT wave[WG / 64 + 1];
  uint32_t v[WG];
  uint32_t hub_n;
  uint32_t hub[WG];
};

// One LLB tile: WG frontier entries (contiguous, or strided by n_tiles for
// SLB), block scan of their degrees, cooperative expansion of their edges.
template <typename PtrT, int WG, bool HOT = false>
__device__ __forceinline__ void llb_tile(const KArgs &a, LlbShared<PtrT, WG> &sh, uint32_t tile,
                                         uint32_t n_tiles, bool strided,
                                         unsigned long long &dacc, bool have_v0 = false,
                                         uint32_t v0 = 0) {
  constexpr int NW = WG / 64;
  PtrT *s_start = sh.start;
  PtrT *s_scan = sh.scan;
  PtrT *s_wave = sh.wave;
  uint32_t *s_v = sh.v;
  uint32_t &s_hub_n = sh.hub_n;
  uint32_t *s_hub = sh.hub;
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  const uint32_t t = threadIdx.x;
  const uint32_t idx = strided ? tile + t * n_tiles : tile * WG + t;
  __syncthreads();  // LDS of the previous tile is no longer read
  if (t == 0) s_hub_n = 0;
  PtrT deg = 0, st = 0;
  uint32_t v = 0;
  if (idx < a.F) {
    v = have_v0 ? v0 : a.q_in[idx];  // v0: this entry, read by the caller
    if (v < a.n_rows) {
      st = col_ptr[v];
      deg = col_ptr[v + 1] - st;
    }
  }
  __syncthreads();
  // Hubs leave the block scan and are expanded wave-per-vertex below.
  if (a.hub_degree && deg >= (PtrT)a.hub_degree) {
    s_hub[atomicAdd(&s_hub_n, 1u)] = t;
  }
  s_start[t] = st;
  s_v[t] = v;
  const PtrT sdeg = (a.hub_degree && deg >= (PtrT)a.hub_degree) ? PtrT(0) : deg;
  PtrT total;
  const PtrT ex = block_excl_scan<WG>(sdeg, total, s_wave);
  s_scan[t] = ex;
  __syncthreads();
  // owner of edge i = last slot j with s_scan[j] <= i (empty slots tie with
  // the next one and are skipped by taking the last)
  expand_edges<PtrT, WG, true, HOT>(a, col_ptr, s_scan, s_start, s_v, PtrT(0), total, dacc);
  if (a.hub_degree) {
    const uint32_t nh = s_hub_n;
    const uint32_t w = t >> 6, l = t & 63;
    for (uint32_t h = w; h < nh; h += NW) {
      const uint32_t slot = s_hub[h];
      const PtrT b = s_start[slot];
      const PtrT e1 = col_ptr[s_v[slot] + 1];
      for (PtrT e = b + l; e < e1; e += 64)
        discover<PtrT, HOT>(a, col_ptr, a.row_idx[e], s_v[slot], dacc);
    }
  }
}

// Host-sized launch: one tile per work-group, grid = ceil(F / WG).  The
// degree sum of exact-once discoveries leaves with one atomic per group.
template <typename PtrT, int WG, bool STRIDED, bool HOT>
__global__ __launch_bounds__(WG) void k_llb(KArgs a) {
  __shared__ LlbShared<PtrT, WG> sh;
  unsigned long long dacc = 0;
  llb_tile<PtrT, WG, HOT>(a, sh, blockIdx.x, gridDim.x, STRIDED, dacc);
  if (a.q_out_edges && !a.mark) block_atomic_add<WG>(dacc, a.q_out_edges);
}

// Device-driven top-down step for captured level loops: the frontier size,
// level, discovery mode and LLB/SLB choice come from the control block, and
// a fixed grid walks the tiles.  Kernels after the search ended see F = 0.
struct DevStep {
  uint32_t cur;                  // queue buffer holding the frontier
  int fixed;                     // BFSX_LLB or BFSX_SLB, or -1: read ctl->choice
  MarkRule mark;                 // which frontiers mark + compact
  int spec;                      // speculative: skip the frontiers `mark` marks
};

static_assert(offsetof(LevelCtl, q_count) == 0 && offsetof(LevelCtl, level) == 8 &&
                  offsetof(LevelCtl, q_edges) == 16 && alignof(LevelCtl) >= 8,
              "k_td_dev reads the control block's first 32 bytes as two uint4");

template <typename PtrT, int WG>
__global__ __launch_bounds__(WG) void k_td_dev(KArgs a, LevelCtl *ctl, DevStep d) {
  __shared__ LlbShared<PtrT, WG> sh;
  // The control block's fields (its first 32 bytes, counts, level and
  // edges, as two 16-byte loads, and the choice) and this group's first LLB
  // tile of the queue in one round trip: the queue holds n_cols + 1
  // entries, so the read is in bounds whatever F is; entries past F are not
  // used.
  const uint4 h0 = reinterpret_cast<const uint4 *>(ctl)[0];
  const uint4 h1 = reinterpret_cast<const uint4 *>(ctl)[1];
  const uint32_t c_choice = ctl->choice;
  const uint32_t idx0 = blockIdx.x * WG + threadIdx.x;
  const uint32_t v0 = idx0 < a.n_cols ? a.q_in[idx0] : 0u;
  const uint32_t F = d.cur ? h0.y : h0.x;
  const unsigned long long c_edges =
      d.cur ? ((unsigned long long)h1.w << 32 | h1.z) : ((unsigned long long)h1.y << 32 | h1.x);
  const int choice = d.fixed >= 0 ? d.fixed : (int)c_choice;
  const bool marks = d.mark.marks(c_edges, F);
  // one branch on all of them, so that no read waits for another's test
  const bool run = (F != 0) & ((choice == BFSX_LLB) | (choice == BFSX_SLB)) &
                   !(marks & (d.spec != 0));
  if (!run) return;
  a.F = F;
  a.next_level = (int32_t)h0.z + 1;
  if (!marks) a.mark = nullptr;
  const uint32_t n_tiles = (F + WG - 1) / WG;
  const bool llb = choice == BFSX_LLB;
  unsigned long long dacc = 0;
  for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    llb_tile<PtrT, WG>(a, sh, tile, n_tiles, !llb, dacc, llb && tile == blockIdx.x, v0);
  if (a.q_out_edges && !a.mark) block_atomic_add<WG>(dacc, a.q_out_edges);
}

// ---------------------------------------------------------------------------
// GLB: tile scan -> tile-offset scan -> global offsets + chunk table -> expand
// ---------------------------------------------------------------------------
constexpr int GLB_TILE = 2048;  // frontier entries per scan tile (256 x 8)
constexpr uint32_t GLB_MIN_CHUNK = 256;  // smallest work-queue item (edges)

template <typename PtrT>
__global__ __launch_bounds__(256) void k_glb_tiles(const PtrT *col_ptr, uint32_t n_rows,
                                                   const uint32_t *q_in, uint32_t F,
                                                   PtrT *g_loc, PtrT *g_start,
                                                   PtrT *tile_sum) {
  __shared__ PtrT s_wave[256 / 64 + 1];
  const uint32_t base = blockIdx.x * GLB_TILE + threadIdx.x * 8;
  PtrT d[8];
  PtrT sum = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t i = base + k;
    PtrT s = 0, dg = 0;
    if (i < F) {
      const uint32_t v = q_in[i];
      if (v < n_rows) {
        s = col_ptr[v];
        dg = col_ptr[v + 1] - s;
      }
      g_start[i] = s;
    }
    d[k] = dg;
    sum += dg;
  }
  PtrT total;
  PtrT run = block_excl_scan<256>(sum, total, s_wave);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t i = base + k;
    if (i < F) g_loc[i] = run;
    run += d[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One work-group turns tile sums into exclusive tile offsets, publishes the
// frontier's edge total and chunk count, and resets the work-queue head.
template <typename PtrT>
__global__ __launch_bounds__(1024) void k_glb_tile_offsets(PtrT *tile_sum, uint32_t n_tiles,
                                                           uint32_t chunk, LevelCtl *ctl,
                                                           PtrT *g_scan_end) {
  __shared__ PtrT s_wave[1024 / 64 + 1];
  __shared__ PtrT s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < n_tiles; b += 4096) {
    PtrT v[4];
    PtrT sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t i = b + threadIdx.x * 4 + k;
      v[k] = i < n_tiles ? tile_sum[i] : PtrT(0);
      sum += v[k];
    }
    PtrT total;
    PtrT run = block_excl_scan<1024>(sum, total, s_wave) + s_carry;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t i = b + threadIdx.x * 4 + k;
      if (i < n_tiles) tile_sum[i] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const PtrT tot = s_carry;
    *g_scan_end = tot;
    ctl->glb_total = (unsigned long long)tot;
    ctl->glb_chunks = (uint32_t)(((unsigned long long)tot + chunk - 1) / chunk);
    ctl->glb_cursor = 0;
  }
}

// Chunk-table entries one thread writes itself; longer ranges (hubs) are
// written by a whole wave or group.
constexpr uint32_t GLB_SMALL_SERIAL = 4;

// ceil(x / chunk): a shift when chunk is a power of two (shift = its log2,
// else 0), which it is for the bucketed passes and the tuned sizes (a 64-bit
// division is a long VALU sequence, and the small scan runs on one CU).
__device__ __forceinline__ unsigned long long chunk_div_up(unsigned long long x, uint32_t chunk,
                                                           uint32_t shift) {
  return shift ? (x + chunk - 1) >> shift : (x + chunk - 1) / chunk;
}

// Global exclusive offsets, and chunk_lo[c] = frontier entry holding edge c*chunk.
template <typename PtrT>
__global__ void k_glb_offsets(const PtrT *g_loc, const PtrT *tile_off, const PtrT *g_scan_end,
                              uint32_t F, PtrT *g_scan, uint32_t chunk, uint32_t shift,
                              uint32_t *chunk_lo) {
  const PtrT total = *g_scan_end;
  const uint32_t lane = threadIdx.x & 63;
  // (every lane runs every round, so that the wave can write long ranges)
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < F; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    unsigned long long c0 = 0, c1 = 0;
    if (i < F) {
      const PtrT ex = g_loc[i] + tile_off[i / GLB_TILE];
      PtrT in;
      if (i + 1 < F)
        in = g_loc[i + 1] + tile_off[(i + 1) / GLB_TILE];
      else
        in = total;
      g_scan[i] = ex;
      // chunks whose first edge falls in [ex, in)
      c0 = chunk_div_up((unsigned long long)ex, chunk, shift);
      c1 = chunk_div_up((unsigned long long)in, chunk, shift);
    }
    const bool longr = c1 - c0 > GLB_SMALL_SERIAL;
    if (!longr)
      for (unsigned long long c = c0; c < c1; c++) chunk_lo[c] = i;
    // a hub's range (hundreds of entries) by the whole wave
    unsigned long long m = __ballot(longr);
    while (m) {
      const int l = __ffsll(m) - 1;
      m &= m - 1;
      const unsigned long long a0 = __shfl(c0, l, 64), a1 = __shfl(c1, l, 64);
      const uint32_t owner = __shfl(i, l, 64);
      for (unsigned long long c = a0 + lane; c < a1; c += 64) chunk_lo[c] = owner;
    }
  }
}

// The three kernels above in one work-group, for frontiers of at most
// GLB_SMALL_F entries (the hub-heavy first levels of R-MAT searches): each
// thread takes 8 consecutive entries per round, a block scan with a running
// carry gives their global offsets, and the chunk table comes from each
// entry's [offset, offset + degree).  Saves two launches and their gaps.
constexpr uint32_t GLB_SMALL_F = 16384;
constexpr int GLB_SMALL_WG = 1024;
constexpr uint32_t GLB_SMALL_BIG = 1024;   // longer ranges listed per round

template <typename PtrT>
__global__ __launch_bounds__(GLB_SMALL_WG) void k_glb_scan_small(
    const PtrT *col_ptr, uint32_t n_rows, const uint32_t *q_in, uint32_t F, PtrT *g_start,
    PtrT *g_scan, uint32_t chunk, uint32_t shift, uint32_t *chunk_lo, LevelCtl *ctl) {
  __shared__ PtrT s_wave[GLB_SMALL_WG / 64 + 1];
  // Entries whose edges span more than GLB_SMALL_SERIAL chunks (hubs) are
  // listed here and their chunk-table ranges written by the whole group: one
  // thread writing a hub's hundreds of entries serialised the kernel.
  __shared__ uint32_t s_nbig;
  __shared__ uint32_t s_big_i[GLB_SMALL_BIG];
  __shared__ unsigned long long s_big_c0[GLB_SMALL_BIG], s_big_c1[GLB_SMALL_BIG];
  if (threadIdx.x == 0) s_nbig = 0;
  __syncthreads();
  PtrT carry = 0;
  for (uint32_t r0 = 0; r0 < F; r0 += GLB_SMALL_WG * 8) {
    const uint32_t base = r0 + threadIdx.x * 8;
    PtrT st[8], d[8];
    PtrT sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t i = base + k;
      st[k] = d[k] = 0;
      if (i < F) {
        const uint32_t v = q_in[i];
        if (v < n_rows) {
          st[k] = col_ptr[v];
          d[k] = col_ptr[v + 1] - st[k];
        }
      }
      sum += d[k];
    }
    PtrT total;
    PtrT ex = block_excl_scan<GLB_SMALL_WG>(sum, total, s_wave) + carry;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t i = base + k;
      if (i < F) {
        g_start[i] = st[k];
        g_scan[i] = ex;
        // chunks whose first edge falls in [ex, ex + d)
        const unsigned long long c0 = chunk_div_up((unsigned long long)ex, chunk, shift);
        const unsigned long long c1 = chunk_div_up((unsigned long long)(ex + d[k]), chunk, shift);
        uint32_t slot = GLB_SMALL_BIG;
        if (c1 - c0 > GLB_SMALL_SERIAL) slot = atomicAdd(&s_nbig, 1u);
        if (slot < GLB_SMALL_BIG) {
          s_big_i[slot] = i;
          s_big_c0[slot] = c0;
          s_big_c1[slot] = c1;
        } else {
          for (unsigned long long c = c0; c < c1; c++) chunk_lo[c] = i;
        }
      }
      ex += d[k];
    }
    carry += total;
    __syncthreads();  // s_wave is rewritten by the next round's scan; the list is complete
    // (a wave per listed entry, its lanes over the entry's range)
    const uint32_t nb = s_nbig < GLB_SMALL_BIG ? s_nbig : GLB_SMALL_BIG;
    for (uint32_t j = threadIdx.x >> 6; j < nb; j += GLB_SMALL_WG / 64)
      for (unsigned long long c = s_big_c0[j] + (threadIdx.x & 63); c < s_big_c1[j]; c += 64)
        chunk_lo[c] = s_big_i[j];
    __syncthreads();
    if (threadIdx.x == 0) s_nbig = 0;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctl->glb_total = (unsigned long long)carry;
    ctl->glb_chunks = (uint32_t)(((unsigned long long)carry + chunk - 1) / chunk);
    ctl->glb_cursor = 0;
  }
}

// STAGE: exact-once discoveries of a chunk (at most chunk <= GLB_STAGE_CAP)
// are staged in LDS and appended with one global atomic per chunk: on a
// level whose edges mostly find new vertices (the first level of an R-MAT
// search), per-wave appends serialise on the one queue counter.
constexpr uint32_t GLB_STAGE_CAP = 2048;

template <typename PtrT, int WG, bool STAGE>
__global__ __launch_bounds__(WG) void k_glb_expand(KArgs a, const PtrT *g_scan,
                                                   const PtrT *g_start,
                                                   const uint32_t *chunk_lo, LevelCtl *ctl,
                                                   uint32_t chunk) {
  __shared__ PtrT s_scan[WG];
  __shared__ PtrT s_start[WG];
  __shared__ uint32_t s_v[WG];
  __shared__ PtrT s_win_lo, s_win_hi;
  __shared__ uint32_t s_c;
  __shared__ uint32_t s_buf[STAGE ? GLB_STAGE_CAP : 1];
  __shared__ uint32_t s_n, s_base;
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  const uint32_t t = threadIdx.x;
  const uint32_t n_chunks = ctl->glb_chunks;
  const PtrT total = (PtrT)ctl->glb_total;
  unsigned long long dacc = 0;
  if (STAGE && t == 0) s_n = 0;
  // the first chunk of each work-group is its own index; later ones come
  // from the cursor (chunks gridDim.x, gridDim.x + 1, ...): a level with no
  // more chunks than work-groups takes no atomic on the one cursor address
  for (bool first = true;; first = false) {
    if (t == 0) s_c = first ? blockIdx.x : gridDim.x + atomicAdd(&ctl->glb_cursor, 1u);
    __syncthreads();
    const uint32_t c = s_c;
    __syncthreads();
    if (c >= n_chunks) break;
    const PtrT e0 = (PtrT)c * chunk;
    const PtrT e1 = (e0 + chunk < total) ? e0 + chunk : total;
    const uint32_t ilo = chunk_lo[c];
    const uint32_t ihi = (c + 1 < n_chunks) ? chunk_lo[c + 1] : a.F - 1;
    for (uint32_t wb = ilo; wb <= ihi; wb += WG) {
      const uint32_t nwin = (ihi + 1 - wb < WG) ? ihi + 1 - wb : WG;
      if (t < nwin) {
        s_scan[t] = g_scan[wb + t];
        s_start[t] = g_start[wb + t];
        if (a.parents) s_v[t] = a.q_in[wb + t];
      } else {
        s_scan[t] = ~PtrT(0);
      }
      if (t == 0) {
        s_win_lo = g_scan[wb];
        s_win_hi = (wb + nwin < a.F) ? g_scan[wb + nwin] : total;
      }
      __syncthreads();
      const PtrT lo = e0 > s_win_lo ? e0 : s_win_lo;
      const PtrT hi = e1 < s_win_hi ? e1 : s_win_hi;
      expand_edges<PtrT, WG, false, false, STAGE>(a, col_ptr, s_scan, s_start, s_v, lo, hi,
                                                   dacc, s_buf, &s_n);
      __syncthreads();
    }
    if (STAGE) {
      // the chunk's discoveries (at most `chunk` of them) to the queue
      if (t == 0) s_base = s_n ? atomicAdd(a.q_out_count, s_n) : 0u;
      __syncthreads();
      for (uint32_t i = t; i < s_n; i += WG) a.q_out[s_base + i] = s_buf[i];
      __syncthreads();
      if (t == 0) s_n = 0;
    }
  }
  if (a.q_out_edges && !a.mark) block_atomic_add<WG>(dacc, a.q_out_edges);
}

// ---------------------------------------------------------------------------
// GLB with bucketed discovery (heavy levels)
// ---------------------------------------------------------------------------
// Random byte stores into `mark` (one per edge whose target is unvisited)
// cost GLB three quarters of its time on heavy levels, and atomics on a
// bitmap cost more.  Bucketed discovery makes every global access sequential
// instead: targets are split by destination slice (2^BKT_SHIFT vertices), so
//   pass 1  each work-group counts its edges per slice in LDS and writes its
//           row of the (group x slice) count matrix;
//   scan    column-wise exclusive scans give each group's slot range inside
//           each slice's segment, and a scan of the slice totals the
//           segments' bases;
//   pass 2  the groups walk the same chunks again (static assignment) and
//           append each target to its slice's segment (LDS cursors), as
//           its 16-bit offset inside the slice;
//   apply   a work-group per slice holds the slice's visited words in LDS,
//           ORs the targets into an LDS found-bitmap, and emits the new
//           vertices: levels, visited and next-frontier words, the queue in
//           vertex order with one atomic per group.
// Every slice fits in LDS (2^16 vertices: 8 KiB visited + 8 KiB found).
constexpr uint32_t BKT_SHIFT = 16;
constexpr uint32_t BKT_MAX = 2048;  // slices: graphs of up to 2^27 vertices
constexpr int BKT_WG = 512;         // work-group size of the two passes
constexpr int BKT_EPT = 4;          // edges per lane per round
constexpr uint32_t BKT_CHUNK = BKT_WG * BKT_EPT;  // edges per chunk: one round

// PASS 1: per-slice counts of this group's edges -> cnt[group][slice].
// PASS 2: slots from pos[group][slice] + base[slice]; targets -> bkt (the
// low 16 bits: the slice is implied by the segment).
// Chunks are assigned statically (group g takes c = g, g + grid, ...), so
// both passes see the same edges in every group.
template <typename PtrT, int WG, int PASS>
__global__ __launch_bounds__(WG) void k_glb_bucket(KArgs a, const PtrT *g_scan,
                                                   const PtrT *g_start,
                                                   const uint32_t *chunk_lo, LevelCtl *ctl,
                                                   uint32_t chunk, uint32_t n_slices,
                                                   uint32_t *cnt, const uint64_t *base,
                                                   uint16_t *bkt, uint64_t cap) {
  __shared__ PtrT s_scan[WG];
  __shared__ PtrT s_start[WG];
  __shared__ PtrT s_win_lo, s_win_hi;
  __shared__ uint32_t s_cnt[BKT_MAX];
  __shared__ uint64_t s_base[PASS == 2 ? BKT_MAX : 1];
  const uint32_t t = threadIdx.x;
  uint32_t *row = cnt + (size_t)blockIdx.x * n_slices;
  for (uint32_t i = t; i < n_slices; i += WG) {
    s_cnt[i] = PASS == 1 ? 0u : row[i];
    if (PASS == 2) s_base[i] = base[i];
  }
  __syncthreads();
  const uint32_t n_chunks = ctl->glb_chunks;
  const PtrT total = (PtrT)ctl->glb_total;
  for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const PtrT e0 = (PtrT)c * chunk;
    const PtrT e1 = (e0 + chunk < total) ? e0 + chunk : total;
    const uint32_t ilo = chunk_lo[c];
    const uint32_t ihi = (c + 1 < n_chunks) ? chunk_lo[c + 1] : a.F - 1;
    for (uint32_t wb = ilo; wb <= ihi; wb += WG) {
      const uint32_t nwin = (ihi + 1 - wb < WG) ? ihi + 1 - wb : WG;
      if (t < nwin) {
        s_scan[t] = g_scan[wb + t];
        s_start[t] = g_start[wb + t];
      } else {
        s_scan[t] = ~PtrT(0);
      }
      if (t == 0) {
        s_win_lo = g_scan[wb];
        s_win_hi = (wb + nwin < a.F) ? g_scan[wb + nwin] : total;
      }
      __syncthreads();
      const PtrT lo = e0 > s_win_lo ? e0 : s_win_lo;
      const PtrT hi = e1 < s_win_hi ? e1 : s_win_hi;
      // Instruction k of a wave takes 64 consecutive edges of the window,
      // one per lane.  Lists are sorted, so inside a list equal slices form
      // runs of adjacent lanes: the first lane of each run updates the LDS
      // counter once for the run (no same-address conflicts), and in pass 2
      // the run's lanes store to consecutive slots.  A run may continue
      // into the next list; its slots are still distinct.
      const bool single = s_scan[1] >= hi;  // the range lies in one list
      if (PASS == 1 && !single) {
        // counting only: EPT consecutive edges per lane and one LDS atomic
        // per edge (measured faster than the runs when lists are short)
        for (PtrT base = lo + (PtrT)t * BKT_EPT; base < hi; base += (PtrT)WG * BKT_EPT) {
          uint32_t j = 0;
#pragma unroll
          for (uint32_t step = WG / 2; step > 0; step >>= 1)
            if (s_scan[j + step] <= base) j += step;
          PtrT idx[BKT_EPT];
#pragma unroll
          for (int k = 0; k < BKT_EPT; k++) {
            const PtrT e = base + k;
            const bool v = e < hi;
            while (v && j + 1 < WG && s_scan[j + 1] <= e) j++;
            // past hi: repeat the previous entry (edge base is valid), so
            // the loads need no branch and are in flight together
            idx[k] = v ? s_start[j] + (e - s_scan[j]) : idx[k > 0 ? k - 1 : 0];
          }
          uint32_t n[BKT_EPT];
#pragma unroll
          for (int k = 0; k < BKT_EPT; k++) n[k] = a.row_idx[idx[k]];
#pragma unroll
          for (int k = 0; k < BKT_EPT; k++)
            if (base + k < hi) atomicAdd(&s_cnt[n[k] >> BKT_SHIFT], 1u);
        }
        __syncthreads();
        continue;
      }
      const uint32_t lane = t & 63;
      const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
      for (PtrT b0 = lo + (PtrT)(t & ~63u) * BKT_EPT; b0 < hi; b0 += (PtrT)WG * BKT_EPT) {
        // owner entry of this lane's first edge (binary search in LDS), then
        // a forward walk for its later edges, 64 apart
        uint32_t j = 0;
        if (!single) {
          const PtrT e = b0 + lane;
#pragma unroll
          for (uint32_t step = WG / 2; step > 0; step >>= 1)
            if (s_scan[j + step] <= e) j += step;
        }
        PtrT idx[BKT_EPT];
#pragma unroll
        for (int k = 0; k < BKT_EPT; k++) {
          const PtrT e = b0 + (PtrT)k * 64 + lane;
          const bool v = e < hi;
          if (!single)
            while (v && j + 1 < WG && s_scan[j + 1] <= e) j++;
          idx[k] = v ? s_start[j] + (e - s_scan[j]) : ~PtrT(0);
        }
        // lanes past hi read lane 0's first entry (always valid: b0 < hi), so
        // the loads need no branch and are in flight together
        const PtrT safe = __shfl(idx[0], 0, 64);
        uint32_t n[BKT_EPT];
#pragma unroll
        for (int k = 0; k < BKT_EPT; k++) n[k] = a.row_idx[idx[k] == ~PtrT(0) ? safe : idx[k]];
#pragma unroll
        for (int k = 0; k < BKT_EPT; k++) {
          const bool v = b0 + (PtrT)k * 64 + lane < hi;
          const uint32_t sl = v ? n[k] >> BKT_SHIFT : 0xFFFFFFFFu;
          const uint32_t prev = __shfl_up(sl, 1, 64);
          const bool head = v && (lane == 0 || prev != sl);
          const unsigned long long hm = __ballot(head);
          if (!hm) continue;
          const unsigned long long vm = __ballot(v);
          uint32_t slot = 0;
          if (head) {
            const unsigned long long later = hm & ~upto;
            const uint32_t end = later ? (uint32_t)(__ffsll(later) - 1) : (uint32_t)__popcll(vm);
            slot = atomicAdd(&s_cnt[sl], end - lane);
          }
          if (PASS == 2) {
            const uint32_t h = 63u - (uint32_t)__clzll(hm & upto);
            const uint32_t hs = __shfl(slot, (int)h, 64);
            if (v) {
              const uint64_t p = s_base[sl] + hs + (lane - h);
              if (p < cap) bkt[p] = (uint16_t)n[k];
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (PASS == 1)
    for (uint32_t i = t; i < n_slices; i += WG) row[i] = s_cnt[i];
}

// Column k of cnt[groups][slices]: exclusive scan down the groups in place;
// the column total goes to tot[k].  A work-group of 16 waves takes 16
// consecutive slices: lane l of a wave reads slice k0 + (l & 15) of row
// (group) 4i + (l >> 4) of its wave's block of rows, so one load instruction
// covers four 64-byte row pieces.  Pass 1 sums each wave's block per slice,
// the block sums are scanned across the waves in LDS, pass 2 rewrites the
// counts as exclusive offsets in row order (the rows just read, from L2).
// (One work-group per slice reading its column, a 64-byte sector per 4-byte
// count, took about 22 us per launch at 1024 x 1024 counts.)
template <int WG>
__global__ __launch_bounds__(WG) void k_bucket_colscan(uint32_t *cnt, uint32_t n_groups,
                                                       uint32_t n_slices, uint64_t *tot) {
  constexpr uint32_t NW = WG / 64;
  __shared__ uint64_t s_part[NW][16];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t sl = lane & 15, j = lane >> 4;
  const uint32_t k = blockIdx.x * 16 + sl;
  const bool live = k < n_slices;
  const uint32_t per = ((n_groups + NW - 1) / NW + 3) & ~3u;  // rows per wave
  const uint32_t g0 = wv * per;
  uint64_t sum = 0;
  for (uint32_t i0 = 0; i0 < per / 4; i0 += 8) {
    // eight loads in flight (an invalid one reads entry 0 and is dropped)
    uint32_t v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
      const uint32_t g = g0 + 4 * (i0 + u) + j;
      const bool ok = live && i0 + u < per / 4 && g < n_groups;
      const uint32_t x = cnt[ok ? (size_t)g * n_slices + k : 0];
      v[u] = ok ? x : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) sum += v[u];
  }
  sum += __shfl_xor(sum, 16, 64);
  sum += __shfl_xor(sum, 32, 64);
  if (j == 0) s_part[wv][sl] = sum;
  __syncthreads();
  if (wv == 0 && j == 0) {
    uint64_t run = 0;
    for (uint32_t w = 0; w < NW; w++) {
      const uint64_t x = s_part[w][sl];
      s_part[w][sl] = run;
      run += x;
    }
    if (live) tot[k] = run;
  }
  __syncthreads();
  uint64_t pos = s_part[wv][sl];
  for (uint32_t i = 0; i < per / 4; i++) {
    const uint32_t g = g0 + 4 * i + j;
    const bool ok = live && g < n_groups;
    const size_t at = (size_t)g * n_slices + k;
    const uint32_t c = ok ? cnt[at] : 0u;
    // rows 4i .. 4i + 3 in order: the counts of the lanes before mine
    const uint32_t c1 = __shfl_up(c, 16, 64), c2 = __shfl_up(c, 32, 64),
                   c3 = __shfl_up(c, 48, 64);
    const uint64_t pre = (uint64_t)(j >= 1 ? c1 : 0u) + (j >= 2 ? c2 : 0u) + (j >= 3 ? c3 : 0u);
    if (ok) cnt[at] = (uint32_t)(pos + pre);  // slice segments hold < 2^32 targets
    uint64_t q = c;
    q += __shfl_xor(q, 16, 64);
    q += __shfl_xor(q, 32, 64);
    pos += q;
  }
}

// The same with a work-group per slice, thread t owning a contiguous run of
// groups: more work-groups where there are few slices (small graphs).
template <int WG>
__global__ __launch_bounds__(WG) void k_bucket_colscan_few(uint32_t *cnt, uint32_t n_groups,
                                                       uint32_t n_slices, uint64_t *tot) {
  __shared__ uint64_t s_wave[WG / 64 + 1];
  const uint32_t k = blockIdx.x;
  const uint32_t run = (n_groups + WG - 1) / WG;
  const uint32_t g0 = threadIdx.x * run;
  const uint32_t g1 = min(g0 + run, n_groups);
  uint64_t sum = 0;
  for (uint32_t g = g0; g < g1; g++) sum += cnt[(size_t)g * n_slices + k];
  uint64_t total;
  uint64_t pos = block_excl_scan<WG>(sum, total, s_wave);
  for (uint32_t g = g0; g < g1; g++) {
    const uint32_t c = cnt[(size_t)g * n_slices + k];
    cnt[(size_t)g * n_slices + k] = (uint32_t)pos;  // slice segments hold < 2^32 targets
    pos += c;
  }
  if (threadIdx.x == 0) tot[k] = total;
}

// base[k] = sum of tot[0..k), base[n_slices] = sum of all (one work-group).
__global__ __launch_bounds__(1024) void k_bucket_base(const uint64_t *tot, uint32_t n_slices,
                                                      uint64_t *base) {
  __shared__ uint64_t s_wave[1024 / 64 + 1];
  const uint32_t t = threadIdx.x;
  const uint32_t run = (n_slices + 1023) / 1024;
  const uint32_t k0 = t * run, k1 = min(k0 + run, n_slices);
  uint64_t sum = 0;
  for (uint32_t k = k0; k < k1; k++) sum += tot[k];
  uint64_t total;
  uint64_t pos = block_excl_scan<1024>(sum, total, s_wave);
  for (uint32_t k = k0; k < k1; k++) {
    base[k] = pos;
    pos += tot[k];
  }
  if (t == 0) base[n_slices] = total;
}

// Bit q = word 2q or 2q + 1 of m is set (64 words -> 32 pairs).
__device__ __forceinline__ uint32_t pair_bits(unsigned long long m) {
  unsigned long long x = (m | (m >> 1)) & 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (uint32_t)x;
}

// A work-group per slice: new vertices of the slice from its targets.
template <typename PtrT, int WG>
__global__ __launch_bounds__(WG) void k_bucket_apply(const PtrT *col_ptr,
                                                     const uint16_t *deg16, uint32_t n_rows,
                                                     uint32_t n_cols, const uint16_t *bkt,
                                                     const uint64_t *base, uint64_t cap,
                                                     uint32_t *vis,
                                                     int32_t *levels, int32_t next_level,
                                                     uint32_t *q_out, uint32_t *q_count,
                                                     unsigned long long *q_edges,
                                                     uint32_t *fbm_next) {
  constexpr uint32_t WPS = 1u << (BKT_SHIFT - 5);  // visited words per slice
  static_assert(WPS % (WG / 64) == 0 && (WPS / (WG / 64)) % 64 == 0, "waves own whole rounds");
  __shared__ uint32_t s_vis[WPS];
  __shared__ uint32_t s_found[WPS];
  __shared__ uint32_t s_qbase;
  const uint32_t k = blockIdx.x;
  const uint32_t n_words = (uint32_t)bitmap_words(n_cols);
  const uint32_t w0 = k * WPS;
  for (uint32_t i = threadIdx.x; i < WPS; i += WG) {
    s_vis[i] = (w0 + i < n_words) ? vis[w0 + i] : ~0u;
    s_found[i] = 0;
  }
  __syncthreads();
  const uint64_t p0 = base[k], p1 = base[k + 1] < cap ? base[k + 1] : cap;
  // eight loads in flight per thread, then the LDS updates
  constexpr uint32_t U = 8;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += (uint64_t)WG * U) {
    uint32_t t[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint64_t q = p + (uint64_t)u * WG;
      const uint32_t x = bkt[q < p1 ? q : p];  // p < p1: no branch around the load
      t[u] = q < p1 ? x : 0xFFFFFFFFu;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t l = t[u];
      if (l == 0xFFFFFFFFu) continue;
      const uint32_t bit = 1u << (l & 31);
      // a plain read skips the atomic for targets already found: at heavy
      // levels many edges lead to the same few vertices (hubs), whose
      // same-word LDS atomics would serialise
      if (!((s_vis[l >> 5] | s_found[l >> 5]) & bit)) atomicOr(&s_found[l >> 5], bit);
    }
  }
  __syncthreads();
  // Output.  Wave w owns words [w * WPW, (w + 1) * WPW) of the slice.
  // Pass A: visited / next-frontier words and the wave's count; one atomic
  // per group reserves the queue slots.  Pass B: 64 consecutive vertices
  // per round, one per lane (coalesced col_ptr reads), slots by ballot and
  // mbcnt, so the queue is in vertex order.
  constexpr uint32_t NW = WG / 64;
  constexpr uint32_t WPW = WPS / NW;  // words per wave
  __shared__ uint32_t s_wcnt[NW + 1];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr uint32_t NJ = WPW / 64;
  unsigned long long wm[NJ];  // the wave's words holding new vertices
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t j = 0; j < NJ; j++) {
    const uint32_t i = wv * WPW + 64 * j + lane;
    const uint32_t w = w0 + i;
    uint32_t nw = 0;
    if (w < n_words) {
      nw = s_found[i] & ~s_vis[i];
      s_found[i] = nw;
      if (nw) vis[w] = s_vis[i] | nw;
      if (fbm_next) fbm_next[w] = nw;
      cnt += __popc(nw);
    }
    wm[j] = __ballot(nw != 0);
  }
  cnt = wave_sum(cnt);
  if (lane == 0) s_wcnt[wv] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < NW; w++) {
      const uint32_t c = s_wcnt[w];
      s_wcnt[w] = run;
      run += c;
    }
    s_qbase = run ? atomicAdd(q_count, run) : 0u;
  }
  __syncthreads();
  // The next frontier's degree sum from the 16-bit degrees (2 bytes per
  // vertex, col_ptr only for saturated ones; two col_ptr entries per vertex
  // made the pass over a dense slice read the whole col_ptr array: 0.5 GB at
  // scale 26).
  uint32_t pos = s_qbase + s_wcnt[wv];
  unsigned long long dsum = 0;
  // Pair P of the wave (bit P of pm) = words 2P and 2P + 1 of its range, set
  // when either holds a new vertex: only those pairs get a round (a round per
  // pair of all 2048 words cost about 40 us of VALU time per launch on levels
  // that find few vertices).  Rounds go eight at a time, their 16-bit degree
  // loads (every lane of the round: one line) in flight together.
  static_assert(NJ <= 4, "two 64-bit pair masks per wave");
  unsigned long long pm[2] = {0ull, 0ull};
#pragma unroll
  for (uint32_t j = 0; j < NJ; j++) pm[j >> 1] |= (unsigned long long)pair_bits(wm[j]) << (32 * (j & 1));
  constexpr uint32_t R = 8;
  while (pm[0] | pm[1]) {
    uint32_t i_[R];
    bool ok[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
      const int h = pm[0] ? 0 : 1;
      ok[r] = (pm[0] | pm[1]) != 0;
      const uint32_t P = ok[r] ? (uint32_t)(64 * h + __ffsll(pm[h]) - 1) : 0u;
      if (ok[r]) pm[h] &= pm[h] - 1;
      i_[r] = wv * WPW + 2 * P;  // (both words below n_words)
    }
    uint32_t d_[R];
    bool take[R];
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
      d_[r] = 0;
      take[r] = false;
      if (!ok[r]) continue;
      const uint32_t i = i_[r];
      const uint32_t word = s_found[i + (lane >> 5)];
      const bool isnew = (word >> (lane & 31)) & 1u;
      const unsigned long long m = __ballot(isnew);
      const uint32_t v = (w0 + i) * 32 + lane;  // below the padded length
      if (isnew) {
        levels[v] = next_level;
        q_out[pos + lanes_below(m)] = v;
      }
      if (q_edges) {
        take[r] = isnew && v < n_rows;
        d_[r] = deg16[v];
      }
      pos += (uint32_t)__popcll(m);
    }
#pragma unroll
    for (uint32_t r = 0; r < R; r++) {
      if (take[r]) {
        const uint32_t v = (w0 + i_[r]) * 32 + lane;
        dsum += d_[r] != DEG16_ESCAPE ? (unsigned long long)d_[r]
                                      : (unsigned long long)(col_ptr[v + 1] - col_ptr[v]);
      }
    }
  }
  if (q_edges) block_atomic_add<WG>(dsum, q_edges);
}

// ---------------------------------------------------------------------------
// PULL (bottom-up)
// ---------------------------------------------------------------------------
// Neighbours a candidate scans in its own lane before the rest of its list
// goes to a quarter wave (a multiple of 4).
constexpr uint32_t PULL_LANE_PREFIX = 256;
// Candidates per lane per round of the bottom-up step.
constexpr int PULL_SLOTS = 2;
// Frontier summary: one bit per 2^shift vertices, at most
// PULL_SUM_WORDS_MAX words, held in dynamic LDS by every work-group of the
// bottom-up step.

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether w is in the current frontier; a clear summary bit answers without
// touching the (L2-missing) frontier bitmap: such a probe reads word 0
// instead, a line every wave keeps hot.  The load is unconditional, so
// consecutive probes have their loads in flight together (a load inside a
// branch is waited for before the branch joins).
// SUM says whether the LDS summary is in use (a template argument, so no
// branch separates the loads of consecutive probes).
template <bool SUM>
__device__ __forceinline__ bool in_frontier(const uint32_t *fbm, const uint32_t *s_sum,
                                            uint32_t shift, uint32_t w) {
  bool may = true;
  if (SUM) {
    const uint32_t b = w >> shift;
    may = (s_sum[b >> 5] >> (b & 31)) & 1u;
  }
  const uint32_t word = fbm[may ? (w >> 5) : 0u];
  return may && ((word >> (w & 31)) & 1u);
}

// Four probes, all four loads issued before any result is used.
template <bool SUM>
__device__ __forceinline__ void in_frontier4(const uint32_t *fbm, const uint32_t *s_sum,
                                             uint32_t shift, const uint32_t (&w)[4],
                                             bool (&h)[4]) {
  bool may[4];
  uint32_t word[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    may[i] = true;
    if (SUM) {
      const uint32_t b = w[i] >> shift;
      may[i] = (s_sum[b >> 5] >> (b & 31)) & 1u;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) word[i] = fbm[may[i] ? (w[i] >> 5) : 0u];
#pragma unroll
  for (int i = 0; i < 4; i++) h[i] = may[i] && ((word[i] >> (w[i] & 31)) & 1u);
}

// Probe of a first neighbour f read from fd: with hub indexing (hbm set) a
// flagged f names hub f & ~PULL_HUB_FLAG, whose frontier bit is in hbm (an
// L2-resident bitmap of the hubs); otherwise f is a vertex id, probed as by
// in_frontier.
template <bool SUM>
__device__ __forceinline__ bool in_frontier_fd(const uint32_t *fbm, const uint32_t *hbm,
                                               const uint32_t *s_sum, uint32_t shift,
                                               uint32_t f) {
  const bool hub = hbm && (f & PULL_HUB_FLAG);
  const uint32_t w = hub ? f & ~PULL_HUB_FLAG : f;
  bool may = true;
  if (SUM && !hub) {
    const uint32_t b = w >> shift;
    may = (s_sum[b >> 5] >> (b & 31)) & 1u;
  }
  const uint32_t *src = hub ? hbm : fbm;
  const uint32_t word = src[may ? (w >> 5) : 0u];
  return may && ((word >> (w & 31)) & 1u);
}

// A step launched ahead of the decision for its level (speculative) runs
// only if that decision, made by the previous level's end, is one of the
// step kinds it was launched for and the frontier is not empty; otherwise
// its kernels, including its level end, exit without a write.
struct SpecGuard {
  const LevelCtl *ctl;  // null: not speculative
  uint32_t consumed;    // the queue counter of the level's frontier
  uint32_t kinds;       // bit k: runs if the decision is kind k
  MarkRule heavy;       // and if this rule does not mark the frontier
  __device__ __forceinline__ bool skip() const {
    if (!ctl) return false;
    const uint32_t F = ctl->q_count[consumed];
    return !((kinds >> (ctl->choice & 31u)) & 1u) || F == 0 ||
           heavy.marks(ctl->q_edges[consumed], F);
  }
};
constexpr MarkRule NEVER_MARKS{~0ull, ~0ull};  // (frontier edges stay below 2^48)

// hbm bit h = frontier bit of hub[h] (H a multiple of 64, and so is the
// grid-stride: every wave runs whole rounds).
__global__ void k_hub_bits(const uint32_t *fbm, const uint32_t *hub, uint32_t H, uint32_t *hbm,
                           SpecGuard guard) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t h0 = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t v0 = h0 < H ? hub[h0] : 0u;  // in flight with the guard's reads
  if (guard.skip()) return;
  for (uint32_t h = h0; h < H; h += stride) {
    const uint32_t v = h == h0 ? v0 : hub[h];
    const unsigned long long m = __ballot((fbm[v >> 5] >> (v & 31)) & 1u);
    if (lane == 0) hbm[h >> 5] = (uint32_t)m;
    if (lane == 32) hbm[h >> 5] = (uint32_t)(m >> 32);
  }
}

// Summary of a frontier bitmap: bit b = any frontier vertex in
// [b << shift, (b + 1) << shift); shift >= 5, so a bit covers whole words.
__global__ void k_fbm_summary(const uint32_t *fbm, uint32_t n_words, uint32_t shift,
                              uint32_t *sum, uint32_t n_bits) {
  const uint32_t per = 1u << (shift - 5);  // bitmap words per summary bit
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n_bits; b0 += stride) {
    const uint32_t b = b0 + threadIdx.x;
    uint32_t any = 0;
    if (b < n_bits) {
      const uint32_t w0 = b * per;
      const uint32_t w1 = min(w0 + per, n_words);
      for (uint32_t w = w0; w < w1; w++) any |= fbm[w];
    }
    const unsigned long long m = __ballot(any != 0);
    const uint32_t lane = threadIdx.x & 63;
    if (lane == 0 && b < n_bits) sum[b >> 5] = (uint32_t)m;
    if (lane == 32 && b < n_bits) sum[b >> 5] = (uint32_t)(m >> 32);
  }
}

// Index of the k-th (0-based) set bit of m, which has more than k set bits:
// the largest p with popc(m & ((1 << p) - 1)) <= k.
__device__ __forceinline__ uint32_t kth_set_bit(uint32_t m, uint32_t k) {
  uint32_t p = 0;
#pragma unroll
  for (uint32_t sh = 16; sh > 0; sh >>= 1)
    if ((uint32_t)__popc(m & ((1u << (p + sh)) - 1u)) <= k) p += sh;
  return p;
}

// Owner word and bit of the idx-th set bit of a wave's 64 words (lane l
// holds word l as m, with excl = set bits in the words before it): the
// last lane whose exclusive count is <= idx (it has a non-zero count, since a
// zero-count lane ties with the next one), by a shuffle binary search.
__device__ __forceinline__ void wave_kth_bit(uint32_t m, uint32_t excl, uint32_t idx,
                                             uint32_t &word, uint32_t &bit) {
  uint32_t o = 0;
#pragma unroll
  for (uint32_t st = 32; st > 0; st >>= 1) {
    const uint32_t ex = __shfl(excl, (int)(o + st), 64);
    if (ex <= idx) o += st;
  }
  const uint32_t om = __shfl(m, (int)o, 64);
  const uint32_t oe = __shfl(excl, (int)o, 64);
  word = o;
  bit = kth_set_bit(om, idx - oe);
}

// The rest of a candidate's list, from entry b to e (wave-collective: every
// lane calls it; cand says whether the lane holds a candidate still to be
// decided).  Every candidate first scans up to PULL_LANE_PREFIX neighbours in
// its own lane, and the rest of a longer list goes to a quarter wave.
template <typename PtrT, bool SUM>
__device__ __forceinline__ void scan_rest(const KArgs &a, const uint32_t *fbm_cur,
                                          const uint32_t *s_sum, uint32_t shift, bool cand,
                                          bool &found, uint32_t &par, PtrT b, const PtrT e,
                                          unsigned long long &scanned) {
  const uint32_t lane = threadIdx.x & 63;
  const bool longer = e - b > PULL_LANE_PREFIX;  // the rest of the list
  const PtrT lim = longer ? b + PULL_LANE_PREFIX : e;
  // Every candidate first scans up to PULL_LANE_PREFIX neighbours in its
  // own lane, four independent probes per round (indices past the end of
  // the prefix repeat its last neighbour, which is harmless): all 64
  // candidates of the wave progress together, and with hub-first lists
  // most stop in the first round.  Longer lists not decided by their
  // prefix continue four at a time, each by a quarter wave (16 lanes x 4
  // loads = 64 consecutive neighbours per round, all loads in flight).
  if (cand && !found) {
    PtrT k = b;
    for (; k < lim; k += 4) {
      const PtrT last = lim - 1;
      const uint32_t w[4] = {a.row_idx[k], a.row_idx[k + 1 < lim ? k + 1 : last],
                             a.row_idx[k + 2 < lim ? k + 2 : last],
                             a.row_idx[k + 3 < lim ? k + 3 : last]};
      bool h[4];
      in_frontier4<SUM>(fbm_cur, s_sum, shift, w, h);
      if (h[0] | h[1] | h[2] | h[3]) {
        found = true;
        par = h[0] ? w[0] : h[1] ? w[1] : h[2] ? w[2] : w[3];
        break;
      }
    }
    if (a.examined) scanned += (unsigned long long)((k + 4 < lim ? k + 4 : lim) - b);
    b = lim;  // where a long list's quarter-wave scan resumes
  }
  unsigned long long big = __ballot(cand && longer && !found);
  const uint32_t q = lane >> 4, ql = lane & 15;
  const unsigned long long qmask = 0xFFFFull << (16 * q);
  while (big) {
    int my_src = -1;  // the candidate lane my quarter serves
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
      if (big) {
        if (i == q) my_src = __ffsll(big) - 1;
        big &= big - 1;
      }
    }
    const PtrT sb = __shfl(b, my_src < 0 ? 0 : my_src, 64);
    const PtrT se = __shfl(e, my_src < 0 ? 0 : my_src, 64);
    PtrT base = my_src < 0 ? se : sb;
    bool done = base >= se, hit = false;
    uint32_t hp = 0;
    while (__ballot(!done)) {
      uint32_t w[4];
      bool h[4];
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) {
        const PtrT k = base + j * 16 + ql;
        // (entry 0 exists: some candidate has a long list)
        w[j] = a.row_idx[(!done && k < se) ? k : PtrT(0)];
      }
      in_frontier4<SUM>(fbm_cur, s_sum, shift, w, h);
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) {
        const PtrT k = base + j * 16 + ql;
        h[j] = h[j] && !done && k < se;
      }
      bool got = false;
      uint32_t gw = 0;
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) {
        const unsigned long long hb = __ballot(h[j]) & qmask;
        const uint32_t pw = __shfl(w[j], hb ? __ffsll(hb) - 1 : (int)lane, 64);
        if (!got && hb) {
          got = true;
          gw = pw;
        }
      }
      if (!done) {
        if (a.examined && ql == 0) scanned += (unsigned long long)(se - base < 64 ? se - base : 64);
        if (got) {
          hit = true;
          hp = gw;
          done = true;
        }
        base += 64;
        if (base >= se) done = true;
      }
    }
    // hand each result to its candidate lane
#pragma unroll
    for (uint32_t qq = 0; qq < 4; qq++) {
      const int s_q = __shfl(my_src, (int)(16 * qq), 64);
      const bool h_q = __shfl((int)hit, (int)(16 * qq), 64) != 0;
      const uint32_t p_q = __shfl(hp, (int)(16 * qq), 64);
      if (s_q == (int)lane) {
        found = h_q;
        par = p_q;
      }
    }
  }
}

// Bottom-up step over plain lists (without the hub-first copy and its
// first-neighbour array; k_pull2 below is the step with them).  A wave owns
// 2048-vertex chunks (grid-stride): lane l
// reads visited word l of the chunk (one coalesced load) and holds its
// unvisited bits.  The unvisited vertices are then checked 64 at a time,
// one per lane: a wave prefix sum of the per-word counts and a shuffle
// binary search give each lane its word, a popcount search its bit (no
// LDS list).  A candidate scans the first PULL_LANE_PREFIX neighbours in
// its lane, the rest of a longer list is scanned by a quarter wave.
// Frontier probes go through the LDS summary when the frontier is
// sparse.  Found vertices set bits in 64 LDS words that the wave then writes
// as the next frontier bitmap and ORs into the visited bitmap (the wave owns
// both word ranges: no global atomics).  The next frontier's size and degree
// sum leave with one atomic per work-group; no queue is written.
template <typename PtrT, int WG, bool SUM>
__global__ __launch_bounds__(WG) void k_pull(KArgs a, const uint32_t *fbm_cur,
                                             uint32_t *fbm_next, const uint32_t *g_sum,
                                             uint32_t sum_words, uint32_t shift,
                                             SpecGuard guard) {
  if (guard.skip()) return;
  constexpr int NW = WG / 64;
  extern __shared__ uint32_t s_sum_mem[];  // sum_words entries when g_sum is set
  __shared__ uint32_t s_found[NW][64];
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t *s_sum = nullptr;
  if (g_sum) {
    for (uint32_t i = threadIdx.x; i < sum_words; i += WG) s_sum_mem[i] = g_sum[i];
    __syncthreads();
    s_sum = s_sum_mem;
  }
  const uint32_t n_words = (uint32_t)bitmap_words(a.n_cols);
  const uint32_t n_chunks = (a.n_rows + PULL_CHUNK - 1) / PULL_CHUNK;
  uint32_t n_found = 0;
  unsigned long long found_deg = 0, scanned = 0;
  uint32_t *fw = s_found[wv];
  for (uint32_t c = blockIdx.x * NW + wv; c < n_chunks; c += gridDim.x * NW) {
    const uint32_t wi = c * 64 + lane;
    const uint32_t vbase = wi * 32;
    const uint32_t vword = wi < n_words ? a.vis[wi] : ~0u;
    uint32_t cm = ~vword;
    if (vbase >= a.n_rows) cm = 0;
    else if (a.n_rows - vbase < 32) cm &= (1u << (a.n_rows - vbase)) - 1u;
    const uint32_t cnt = __popc(cm);
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - cnt;
    fw[lane] = 0;
    wave_sync_lds();
    for (uint32_t r = 0; r < total; r += 64 * PULL_SLOTS) {
      // PULL_SLOTS candidates per lane per round: their col_ptr loads are in
      // flight together; the scans follow slot by slot, uniformly over the
      // wave.
      bool cand_[PULL_SLOTS];
      uint32_t off_[PULL_SLOTS], u_[PULL_SLOTS];
      PtrT b_[PULL_SLOTS], e_[PULL_SLOTS], deg_[PULL_SLOTS];
#pragma unroll
      for (int sl = 0; sl < PULL_SLOTS; sl++) {
        const uint32_t idx = r + 64 * sl + lane;
        cand_[sl] = idx < total;
        uint32_t o, bit;
        wave_kth_bit(cm, excl, idx, o, bit);
        off_[sl] = cand_[sl] ? o * 32 + bit : 0u;
        u_[sl] = c * PULL_CHUNK + off_[sl];
        b_[sl] = e_[sl] = deg_[sl] = 0;
      }
#pragma unroll
        for (int sl = 0; sl < PULL_SLOTS; sl++) {
          const PtrT cb = col_ptr[u_[sl]], ce = col_ptr[u_[sl] + 1];
          if (cand_[sl]) {
            b_[sl] = cb;
            e_[sl] = ce;
            deg_[sl] = ce - cb;
          }
        }
#pragma unroll
      for (int sl = 0; sl < PULL_SLOTS; sl++) {
        const bool cand = cand_[sl];
        const uint32_t off = off_[sl], u = u_[sl];
        bool found = false;
        uint32_t par = 0;
        const PtrT b = b_[sl], e = e_[sl], deg = deg_[sl];
        scan_rest<PtrT, SUM>(a, fbm_cur, s_sum, shift, cand, found, par, b, e, scanned);
        if (found) {
          a.levels[u] = a.next_level;
          if (a.parents) a.parents[u] = par;
          atomicOr(&fw[off >> 5], 1u << (off & 31));
          found_deg += (unsigned long long)deg;
        }
      }  // slots
    }
    wave_sync_lds();
    const uint32_t f = fw[lane];
    if (wi < n_words) {
      fbm_next[wi] = f;
      if (f) a.vis[wi] = vword | f;
    }
    n_found += __popc(f);
    wave_sync_lds();  // fw is cleared by the next chunk
  }
  // one atomic per work-group for each total
  __shared__ uint32_t s_n[NW];
  __shared__ unsigned long long s_d[NW], s_x[NW];
  const uint32_t nf = wave_sum(n_found);
  const unsigned long long dg = wave_sum(found_deg);
  const unsigned long long sx = a.examined ? wave_sum(scanned) : 0ull;
  if (lane == 0) {
    s_n[wv] = nf;
    s_d[wv] = dg;
    s_x[wv] = sx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0;
    unsigned long long d = 0, x = 0;
    for (int w = 0; w < NW; w++) {
      n += s_n[w];
      d += s_d[w];
      x += s_x[w];
    }
    if (n) atomicAdd(a.q_out_count, n);
    if (a.q_out_edges && d) atomicAdd(a.q_out_edges, d);
    if (a.examined && x) atomicAdd(a.examined, x);
  }
}


// Bottom-up step with the hub-first copy and the first-neighbour array, in
// two phases per chunk.  Phase 1 decides every candidate it can from its
// first (highest-degree) neighbour alone: fd entry, one probe (of the hub
// bitmap for a hub); candidates left undecided with more neighbours get a
// bit in an LDS mask.  Phase 2 takes those 64 * PULL_SLOTS at a time: their
// col_ptr pair, then the rest of the list (scan_rest).
// Separating them keeps the short rounds of phase 1 from waiting for the
// long chain (col_ptr, list line, probes) of a few undecided lanes.
// Otherwise as k_pull: found vertices set bits in 64 LDS words the wave then
// writes as the next frontier bitmap and ORs into the visited bitmap; sizes
// and degree sums leave with one atomic per work-group.
template <typename PtrT, int WG, bool SUM>
__global__ __launch_bounds__(WG) void k_pull2(KArgs a, const uint32_t *fbm_cur,
                                              uint32_t *fbm_next, const uint32_t *g_sum,
                                              uint32_t sum_words, uint32_t shift,
                                              const uint32_t *fd, const uint16_t *fdd,
                                              const uint32_t *iso,
                                              const uint32_t *crank, const uint32_t *hbm,
                                              const uint32_t *hub, SpecGuard guard) {
  if (guard.skip()) return;
  constexpr int NW = WG / 64;
  constexpr int S1 = 2 * PULL_SLOTS;  // phase-1 candidates per lane per round
  extern __shared__ uint32_t s_sum_mem[];  // sum_words entries when g_sum is set
  __shared__ uint32_t s_found[NW][64];
  __shared__ uint32_t s_undec[NW][64];
  const PtrT *col_ptr = static_cast<const PtrT *>(a.col_ptr);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t *s_sum = nullptr;
  if (g_sum) {
    for (uint32_t i = threadIdx.x; i < sum_words; i += WG) s_sum_mem[i] = g_sum[i];
    __syncthreads();
    s_sum = s_sum_mem;
  }
  const uint32_t n_words = (uint32_t)bitmap_words(a.n_cols);
  const uint32_t n_chunks = (a.n_rows + PULL_CHUNK - 1) / PULL_CHUNK;
  uint32_t n_found = 0;
  unsigned long long found_deg = 0, scanned = 0;
  uint32_t *fw = s_found[wv];
  uint32_t *uw = s_undec[wv];
  for (uint32_t c = blockIdx.x * NW + wv; c < n_chunks; c += gridDim.x * NW) {
    const uint32_t wi = c * 64 + lane;
    const uint32_t vbase = wi * 32;
    const uint32_t vword = wi < n_words ? a.vis[wi] : ~0u;
    uint32_t cm = ~vword;
    uint32_t lw = wi < n_words ? ~iso[wi] : 0u;  // vertices with neighbours
    if (vbase >= a.n_rows) {
      cm = 0;
      lw = 0;
    } else if (a.n_rows - vbase < 32) {
      cm &= (1u << (a.n_rows - vbase)) - 1u;
      lw &= (1u << (a.n_rows - vbase)) - 1u;
    }
    const uint32_t cnt = __popc(cm);
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - cnt;
    fw[lane] = 0;
    uw[lane] = 0;
    wave_sync_lds();
    if (total) {
      // fd entry of a vertex of this chunk: its rank among the vertices with
      // neighbours (the candidates are among them: isolated ones start
      // visited)
      const uint32_t lc = __popc(lw);
      const uint32_t lexcl = wave_incl_scan(lc) - lc;
      const uint32_t lbase = crank[c];
      // (its shuffles run on every lane: a shuffle from a lane that is not
      // active returns 0, so callers select afterwards)
      auto rank_of = [&](uint32_t o, uint32_t bit) {
        const uint32_t ol = __shfl(lw, (int)o, 64);
        const uint32_t ole = __shfl(lexcl, (int)o, 64);
        return lbase + ole + __popc(ol & ((1u << bit) - 1u));
      };
      // phase 1: first neighbours
      for (uint32_t r = 0; r < total; r += 64 * S1) {
        bool cand[S1];
        uint32_t off[S1], fi[S1];
#pragma unroll
        for (int sl = 0; sl < S1; sl++) {
          const uint32_t idx = r + 64 * sl + lane;
          cand[sl] = idx < total;
          uint32_t o = 0, bit = 0, rk = lbase;
          if (r + 64 * sl < total) {  // (uniform: slots past the chunk's last do nothing)
            wave_kth_bit(cm, excl, idx, o, bit);
            rk = rank_of(o, bit);
          }
          off[sl] = cand[sl] ? o * 32 + bit : 0u;
          // (not a candidate: the chunk's first vertex with neighbours)
          fi[sl] = cand[sl] ? rk : lbase;
        }
        uint32_t f[S1], d16[S1];
#pragma unroll
        for (int sl = 0; sl < S1; sl++) {
          f[sl] = fd[fi[sl]];
          d16[sl] = fdd[fi[sl]];
        }
        bool hf[S1];
#pragma unroll
        for (int sl = 0; sl < S1; sl++)
          hf[sl] = in_frontier_fd<SUM>(fbm_cur, hbm, s_sum, shift, cand[sl] ? f[sl] : 0u);
#pragma unroll
        for (int sl = 0; sl < S1; sl++) {
          if (!cand[sl]) continue;
          const uint32_t u = c * PULL_CHUNK + off[sl];
          if (a.examined) scanned += 1;
          if (hf[sl]) {
            a.levels[u] = a.next_level;
            if (a.parents)
              a.parents[u] = (hbm && (f[sl] & PULL_HUB_FLAG)) ? hub[f[sl] & ~PULL_HUB_FLAG] : f[sl];
            atomicOr(&fw[off[sl] >> 5], 1u << (off[sl] & 31));
            found_deg += d16[sl] != DEG16_ESCAPE ? (unsigned long long)d16[sl]
                                                 : (unsigned long long)(col_ptr[u + 1] - col_ptr[u]);
          } else if (d16[sl] > 1) {
            atomicOr(&uw[off[sl] >> 5], 1u << (off[sl] & 31));
          }
        }
      }
      wave_sync_lds();
      // phase 2: the rest of the undecided candidates' lists
      const uint32_t um = uw[lane];
      const uint32_t cnt2 = __popc(um);
      const uint32_t incl2 = wave_incl_scan(cnt2);
      const uint32_t total2 = __shfl(incl2, 63, 64);
      const uint32_t excl2 = incl2 - cnt2;
      for (uint32_t r = 0; r < total2; r += 64 * PULL_SLOTS) {
        bool cand[PULL_SLOTS];
        uint32_t off[PULL_SLOTS];
#pragma unroll
        for (int sl = 0; sl < PULL_SLOTS; sl++) {
          const uint32_t idx = r + 64 * sl + lane;
          cand[sl] = idx < total2;
          uint32_t o = 0, bit = 0;
          if (r + 64 * sl < total2) wave_kth_bit(um, excl2, idx, o, bit);
          off[sl] = cand[sl] ? o * 32 + bit : 0u;
        }
        // (u is a vertex of this chunk below n_rows for every lane)
        PtrT cb[PULL_SLOTS], ce[PULL_SLOTS];
#pragma unroll
        for (int sl = 0; sl < PULL_SLOTS; sl++) {
          cb[sl] = col_ptr[c * PULL_CHUNK + off[sl]];
          ce[sl] = col_ptr[c * PULL_CHUNK + off[sl] + 1];
        }
#pragma unroll
        for (int sl = 0; sl < PULL_SLOTS; sl++) {
          const PtrT deg = cand[sl] ? ce[sl] - cb[sl] : PtrT(0);
          bool found = false;
          uint32_t par = 0;
          // (the first neighbour is done; not a candidate: an empty range)
          scan_rest<PtrT, SUM>(a, fbm_cur, s_sum, shift, cand[sl], found, par, cb[sl] + 1,
                               cand[sl] ? cb[sl] + deg : cb[sl] + 1, scanned);
          if (found) {
            const uint32_t u = c * PULL_CHUNK + off[sl];
            a.levels[u] = a.next_level;
            if (a.parents) a.parents[u] = par;
            atomicOr(&fw[off[sl] >> 5], 1u << (off[sl] & 31));
            found_deg += (unsigned long long)deg;
          }
        }
      }
    }
    wave_sync_lds();
    const uint32_t f = fw[lane];
    if (wi < n_words) {
      fbm_next[wi] = f;
      if (f) a.vis[wi] = vword | f;
    }
    n_found += __popc(f);
    wave_sync_lds();  // fw and uw are cleared by the next chunk
  }
  // one atomic per work-group for each total
  __shared__ uint32_t s_n[NW];
  __shared__ unsigned long long s_d[NW], s_x[NW];
  const uint32_t nf = wave_sum(n_found);
  const unsigned long long dg = wave_sum(found_deg);
  const unsigned long long sx = a.examined ? wave_sum(scanned) : 0ull;
  if (lane == 0) {
    s_n[wv] = nf;
    s_d[wv] = dg;
    s_x[wv] = sx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0;
    unsigned long long d = 0, x = 0;
    for (int w = 0; w < NW; w++) {
      n += s_n[w];
      d += s_d[w];
      x += s_x[w];
    }
    if (n) atomicAdd(a.q_out_count, n);
    if (a.q_out_edges && d) atomicAdd(a.q_out_edges, d);
    if (a.examined && x) atomicAdd(a.examined, x);
  }
}

// Frontier queue from a frontier bitmap (after a bottom-up step, when a
// top-down step follows).  A work-group owns words_per_group consecutive
// words (a multiple of WG) and each thread a contiguous run of them, so the
// queue comes out in vertex order.  Pass 1 counts, one atomicAdd reserves
// the group's slots, pass 2 writes.
template <int WG>
__global__ __launch_bounds__(WG) void k_bm_to_queue(const uint32_t *bm, uint32_t n_words,
                                                   uint32_t words_per_group, uint32_t *q,
                                                   uint32_t *counter) {
  __shared__ uint32_t s_wave[WG / 64 + 1];
  __shared__ uint32_t s_base;
  const uint32_t run = words_per_group / WG;
  const uint32_t w0 = blockIdx.x * words_per_group + threadIdx.x * run;
  const uint32_t w1 = min(w0 + run, n_words);
  uint32_t cnt = 0;
  for (uint32_t w = w0; w < w1; w++) cnt += __popc(bm[w]);
  uint32_t total;
  const uint32_t ex = block_excl_scan<WG>(cnt, total, s_wave);
  if (threadIdx.x == 0) s_base = total ? atomicAdd(counter, total) : 0u;
  __syncthreads();
  uint32_t pos = s_base + ex;
  for (uint32_t w = w0; w < w1; w++)
    for (uint32_t b = bm[w]; b; b &= b - 1) q[pos++] = w * 32 + (__ffs(b) - 1);
}

// ---------------------------------------------------------------------------
// level generation after a mark step (the reference's LevelGen, :481-521)
// ---------------------------------------------------------------------------
// 4-bit mask of the non-zero bytes of w.
__device__ __forceinline__ uint32_t nz_bytes(uint32_t w) {
  uint32_t t = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
  return ((t >> 7) & 1u) | ((t >> 14) & 2u) | ((t >> 21) & 4u) | ((t >> 28) & 8u);
}

// A work-group owns a run of up to COMPACT_MAX_CHUNKS consecutive
// 1024-vertex chunks; a wave works on one chunk at a time, each lane reading
// 16 mark bytes as one uint4.  Pass 1 counts the marked vertices of every
// chunk into LDS; one wave scans the counts and reserves the group's queue
// slots with a single atomicAdd (so the queue stays in vertex order inside a
// group).  Pass 2 re-reads the marks (now in L2), gives new vertices their
// level, writes them to the queue, ORs their visited bits into words this
// wave owns (no atomics), writes the next frontier bitmap words in full and
// clears the marks.  The degree sum goes out with one atomicAdd per group.
constexpr uint32_t COMPACT_MAX_CHUNKS = 64;
constexpr uint32_t COMPACT_WG = 256;

__device__ __forceinline__ uint32_t mark_bits(uint4 m) {
  if (!(m.x | m.y | m.z | m.w)) return 0u;
  return nz_bytes(m.x) | (nz_bytes(m.y) << 4) | (nz_bytes(m.z) << 8) | (nz_bytes(m.w) << 12);
}

template <typename PtrT>
__device__ __forceinline__ void compact_body(const PtrT *col_ptr, const uint16_t *deg16,
                                             uint32_t n_rows, uint8_t *mark,
                                             uint32_t n_chunks, uint32_t chunks_per_group,
                                             uint32_t *vis, int32_t *levels, int32_t next_level,
                                             uint32_t *q_out, uint32_t *q_count,
                                             unsigned long long *q_edges, uint32_t *fbm_next) {
  constexpr uint32_t NW = COMPACT_WG / 64;
  __shared__ uint32_t s_num[COMPACT_MAX_CHUNKS];
  __shared__ uint32_t s_pre[COMPACT_MAX_CHUNKS];
  __shared__ uint32_t s_base;
  __shared__ unsigned long long s_deg[NW];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t c0 = blockIdx.x * chunks_per_group;
  if (c0 >= n_chunks) return;  // uniform over the group
  const uint32_t nc = min(chunks_per_group, n_chunks - c0);
  uint4 *mk = reinterpret_cast<uint4 *>(mark);
  for (uint32_t i = wv; i < nc; i += NW) {
    const uint32_t t = wave_sum((uint32_t)__popc(mark_bits(mk[(size_t)(c0 + i) * 64 + lane])));
    if (lane == 0) s_num[i] = t;
  }
  __syncthreads();
  if (wv == 0) {
    const uint32_t n = lane < nc ? s_num[lane] : 0u;
    const uint32_t incl = wave_incl_scan(n);
    const uint32_t total = __shfl(incl, 63, 64);
    if (lane < nc) s_pre[lane] = incl - n;
    if (lane == 0) s_base = total ? atomicAdd(q_count, total) : 0u;
  }
  __syncthreads();
  unsigned long long dsum = 0;
  for (uint32_t i = wv; i < nc; i += NW) {
    const uint32_t c = c0 + i;
    uint32_t bits = 0;
    if (s_num[i]) {
      uint4 *mp = mk + (size_t)c * 64 + lane;
      bits = mark_bits(*mp);
      if (bits) *mp = make_uint4(0, 0, 0, 0);
      const uint32_t cnt = __popc(bits);
      uint32_t pos = s_base + s_pre[i] + wave_incl_scan(cnt) - cnt;
      const uint32_t v0 = c * 1024 + lane * 16;
      for (uint32_t b = bits; b; b &= b - 1) {
        const uint32_t v = v0 + (__ffs(b) - 1);
        levels[v] = next_level;
        q_out[pos++] = v;
        if (q_edges && v < n_rows) dsum += degree16(deg16, col_ptr, v);
      }
    }
    const uint32_t partner = __shfl_xor(bits, 1, 64);
    if ((lane & 1) == 0) {
      const uint32_t word = bits | (partner << 16);
      const uint32_t wi = c * 32 + (lane >> 1);
      if (word) vis[wi] |= word;
      if (fbm_next) fbm_next[wi] = word;
    }
  }
  if (q_edges) {
    dsum = wave_sum(dsum);
    if (lane == 0) s_deg[wv] = dsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long d = 0;
      for (uint32_t w = 0; w < NW; w++) d += s_deg[w];
      if (d) atomicAdd(q_edges, d);
    }
  }
}

template <typename PtrT>
__global__ __launch_bounds__(COMPACT_WG) void k_compact(const PtrT *col_ptr,
                                                        const uint16_t *deg16, uint32_t n_rows,
                                                        uint8_t *mark, uint32_t n_chunks,
                                                        uint32_t chunks_per_group, uint32_t *vis,
                                                        int32_t *levels, int32_t next_level,
                                                        uint32_t *q_out, uint32_t *q_count,
                                                        unsigned long long *q_edges,
                                                        uint32_t *fbm_next) {
  compact_body<PtrT>(col_ptr, deg16, n_rows, mark, n_chunks, chunks_per_group, vis, levels,
                     next_level, q_out, q_count, q_edges, fbm_next);
}

// Compaction inside a captured level loop: runs only when the step of this
// level marked (same rule as k_td_dev), level from the control block.
template <typename PtrT>
__global__ __launch_bounds__(COMPACT_WG) void k_compact_dev(
    const PtrT *col_ptr, const uint16_t *deg16, uint32_t n_rows, uint8_t *mark, uint32_t n_chunks,
    uint32_t chunks_per_group, uint32_t *vis, int32_t *levels, LevelCtl *ctl, DevStep d,
    uint32_t *q_out, int track, uint32_t *fbm_next) {
  const uint32_t F = ctl->q_count[d.cur];
  if (F == 0 || !d.mark.marks(ctl->q_edges[d.cur], F)) return;
  const uint32_t out = d.cur ^ 1;
  compact_body<PtrT>(col_ptr, deg16, n_rows, mark, n_chunks, chunks_per_group, vis, levels,
                     (int32_t)ctl->level + 1, q_out, &ctl->q_count[out],
                     track ? &ctl->q_edges[out] : nullptr, fbm_next);
}

// frontier bitmap from a frontier queue: clear, then set one bit per entry
__global__ void k_fill_words(uint32_t *p, uint32_t n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

__global__ void k_queue_bits(const uint32_t *q, uint32_t F, uint32_t *fbm) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < F; i += stride) {
    const uint32_t v = q[i];
    atomicOr(&fbm[v >> 5], 1u << (v & 31));
  }
}


// ---------------------------------------------------------------------------
// search set-up and per-level bookkeeping
// ---------------------------------------------------------------------------
template <typename PtrT>
__global__ void k_init(const PtrT *col_ptr, uint32_t n_rows, uint32_t n_cols, uint32_t root,
                       int32_t *levels, uint32_t *parents, uint32_t *vis,
                       const uint32_t *iso_bm, uint32_t n_words, uint32_t *queue0,
                       LevelCtl *ctl) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cols; i += stride) {
    levels[i] = (i == root) ? 0 : -1;
    if (parents) parents[i] = (i == root) ? root : 0xFFFFFFFFu;
  }
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n_words; w += stride) {
    uint32_t x = iso_bm ? iso_bm[w] : 0u;
    if (w == (root >> 5)) x |= 1u << (root & 31);
    vis[w] = x;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    queue0[0] = root;
    const unsigned long long d =
        root < n_rows ? (unsigned long long)(col_ptr[root + 1] - col_ptr[root]) : 0ull;
    LevelCtl c{};
    c.q_count[0] = 1;
    c.q_count[1] = 0;
    c.level = 0;
    c.q_edges[0] = d;
    c.visited_vertices = 1;
    c.visited_edges = d;
    c.choice = BFSX_LLB;
    c.seq = ctl->seq;  // the level-end counter runs on across searches
    *ctl = c;
  }
}

__global__ void k_level_begin(LevelCtl *ctl) {
  if (threadIdx.x == 0) ctl->t_begin = (unsigned long long)wall_clock64();
}

// The values of the host's level-end words (see write_level_end).
struct LevelEndWords {
  uint32_t F, choice, done, level, seq;
  unsigned long long EF, VE;
};

// Bookkeeping at the end of a level, by one work-group (every thread calls
// it; the scheduler's inference uses the first 64).  kind >= 0: the step
// kind that ran; kind < 0: the one in ctl->choice.  A level with an empty
// frontier (after the end, in a captured loop) changes nothing but the done
// flag.  Returns false, having changed nothing, when `guard` says that a
// speculative launch's guess failed.  With `words`, thread 0 also leaves
// there what the control block holds afterwards for the host's words, so
// that they need no second read of it.
__device__ __forceinline__ bool level_end_body(LevelCtl *ctl, uint32_t consumed,
                                               uint32_t produced, int stamp, int kind,
                                               const float *Wg, uint64_t n_cols, uint64_t nnz,
                                               uint64_t n_iso, uint32_t allow, int reuse,
                                               const SpecGuard &guard,
                                               LevelEndWords *words = nullptr) {
  __shared__ int s_live, s_skip;
  __shared__ float4 s_W4[NPARAM / 4];
  // the network's weights, loaded while thread 0 reads the control block
  // (both are round trips to memory; one waits for the other otherwise)
  constexpr int WPT = (NPARAM / 4 + 63) / 64;
  float4 wr[WPT];
  if (Wg && threadIdx.x < 64) {
    const float4 *G4 = reinterpret_cast<const float4 *>(Wg);
#pragma unroll
    for (int i = 0; i < WPT; i++) {
      const int j = (int)threadIdx.x + 64 * i;
      wr[i] = G4[j < NPARAM / 4 ? j : 0];
    }
  }
  if (threadIdx.x == 0) {
    // every field the bookkeeping reads, loaded together before any store
    const uint32_t fc = ctl->q_count[consumed];
    const uint32_t n = ctl->q_count[produced];
    const uint32_t choice = ctl->choice;
    const unsigned long long ef = ctl->q_edges[produced];
    const unsigned long long vv = ctl->visited_vertices, ve = ctl->visited_edges;
    const uint32_t level = ctl->level;
    const uint32_t kc0 = ctl->kind_count[0], kc1 = ctl->kind_count[1],
                   kc2 = ctl->kind_count[2], kc3 = ctl->kind_count[3];
    const uint32_t f0 = ctl->dec_F, d_pull = ctl->dec_pull, d_valid = ctl->dec_valid;
    const unsigned long long e0 = ctl->dec_EF;
    const unsigned long long ec = ctl->q_edges[consumed];
    const uint32_t seq0 = ctl->seq;
    // a speculative launch whose guess failed changes nothing
    const bool skip = guard.ctl && (!((guard.kinds >> (choice & 31u)) & 1u) || fc == 0 ||
                                    guard.heavy.marks(ec, fc));
    s_skip = skip;
    s_live = 0;
    if (skip) {
      // nothing to do
    } else if (fc == 0) {
      ctl->done = 1;
      if (words) *words = LevelEndWords{n, choice, 1u, level, seq0, ef, ve};
    } else {
      if (words) *words = LevelEndWords{n, choice, n == 0 ? 1u : 0u, level + 1, seq0, ef, ve + ef};
      if (stamp) ctl->t_end = (unsigned long long)wall_clock64();
      const uint32_t k = kind >= 0 ? (uint32_t)kind : choice;
      if (k < 4) ctl->kind_count[k] = (k == 0 ? kc0 : k == 1 ? kc1 : k == 2 ? kc2 : kc3) + 1;
      ctl->discovered = n;
      ctl->visited_vertices = vv + n;
      ctl->visited_edges = ve + ef;
      ctl->q_count[consumed] = 0;
      ctl->q_edges[consumed] = 0;
      ctl->glb_cursor = 0;
      ctl->conv = 0;
      ctl->level = level + 1;
      ctl->done = n == 0;
      s_live = n != 0;
      if (s_live && Wg) {
        // keep the last decision while the frontier (vertices and edges)
        // stays within a factor 1.5 of the one it was computed for, after
        // the same kind of step
        const uint32_t pull = k == BFSX_PULL;
        const bool close = (uint64_t)n * 3 >= (uint64_t)f0 * 2 && (uint64_t)n * 2 <= (uint64_t)f0 * 3 &&
                           ef * 3 >= e0 * 2 && ef * 2 <= e0 * 3;
        if (reuse && d_valid && d_pull == pull && close) {
          s_live = 0;
        } else {
          ctl->dec_F = n;
          ctl->dec_EF = ef;
          ctl->dec_pull = pull;
          ctl->dec_valid = 1;
        }
      }
    }
  }
  __syncthreads();
  if (s_skip) return false;
  if (Wg) {  // uniform over the work-group
    // the learned scheduler picks the next level's step in the same launch
    // (a level that keeps the last decision skips the network)
    if (s_live) {
      if (threadIdx.x < 64) {
#pragma unroll
        for (int i = 0; i < WPT; i++) {
          const int j = (int)threadIdx.x + 64 * i;
          if (j < NPARAM / 4) s_W4[j] = wr[i];
        }
      }
      __syncthreads();
      const uint32_t picked = decide_compute(ctl, s_W4, produced, n_cols, nnz, n_iso, allow, 0);
      if (words && threadIdx.x == 0) words->choice = picked;
    }
  }
  return true;
}

// What the host needs for the next level, into the host-mapped words, each
// tagged with a new sequence number (by the thread that made, or saw, every
// write to the control block): the host polls them instead of synchronising
// the stream.  `produced` = the queue buffer holding the next frontier.
// No fence follows: the host checks every word by its own tag (device
// writes to host memory can arrive in any order anyway), and a system-scope
// fence here only kept the kernel, and the stream behind it, waiting for
// the host's acknowledgement.
__device__ __forceinline__ void put_level_end(LevelCtl *ctl, const LevelEndWords &v,
                                              LevelMirror *mirror) {
  const uint32_t seq = v.seq + 1;
  ctl->seq = seq;
  const unsigned long long t32 = (unsigned long long)seq << 32;
  const unsigned long long t16 = (unsigned long long)(seq & 0xFFFFu) << 48;
  mirror += seq & 1u;  // the set for this parity
  mirror->w[0] = t32 | v.F;
  mirror->w[1] = t16 | (v.EF & MIRROR_48);
  mirror->w[2] = t16 | (v.VE & MIRROR_48);
  mirror->w[3] = t32 | (v.choice & 0xFFu) | ((unsigned long long)(v.done ? 1u : 0u) << 8) |
                 ((unsigned long long)(v.level & MIRROR_LEVEL_MASK) << 9);
}

__device__ __forceinline__ void write_level_end(LevelCtl *ctl, uint32_t produced,
                                                LevelMirror *mirror) {
  put_level_end(ctl,
                LevelEndWords{ctl->q_count[produced], ctl->choice, ctl->done, ctl->level,
                              ctl->seq, ctl->q_edges[produced], ctl->visited_edges},
                mirror);
}

__global__ __launch_bounds__(64) void k_level_end(LevelCtl *ctl, uint32_t consumed,
                                                  uint32_t produced, int stamp, int kind,
                                                  const float *Wg, uint64_t n_cols, uint64_t nnz,
                                                  uint64_t n_iso, uint32_t allow, int reuse,
                                                  LevelMirror *mirror, SpecGuard guard) {
  LevelEndWords words;
  if (!level_end_body(ctl, consumed, produced, stamp, kind, Wg, n_cols, nnz, n_iso, allow, reuse,
                      guard, &words))
    return;
  if (mirror && threadIdx.x == 0) put_level_end(ctl, words, mirror);
}

// ---------------------------------------------------------------------------
// device loop: consecutive thin top-down levels in one cooperative launch
// ---------------------------------------------------------------------------
// Every work-group is resident (cooperative launch), so levels are separated
// by grid-wide barriers instead of kernel launches and host round trips.  A
// level runs while it does not mark (exact-once discovery, the regime where
// a level costs microseconds) and its kind is
// LLB or SLB (the fixed kind, or the scheduler's choice, made at each level
// end, inside the barrier, by the last work-group to arrive, exactly as
// k_level_end makes it).  The loop leaves at
// the first level that is heavy, or for which the scheduler picks GLB or
// PULL, or when the frontier is empty, or after max_levels; work-group 0
// then writes the tagged level-end words (with the level reached) for the
// host, which runs the next level itself.
struct CoopArgs {
  uint32_t *queue[2];
  uint32_t cur;                  // queue buffer holding the frontier at entry
  int fixed;                     // BFSX_LLB or BFSX_SLB, or -1: ctl->choice
  MarkRule mark;                 // frontiers it marks leave the loop
  uint32_t max_levels;
  const float *Wg;               // scheduler weights (learned) or null
  uint64_t n_cols, nnz, n_iso;
  uint32_t allow;                // kinds the scheduler may pick
  LevelMirror *mirror;
};

// Grid-wide barrier of a cooperative launch (every work-group is resident),
// with the level's bookkeeping inside it: one arrival atomic per
// work-group; the last group to arrive runs `body` (all its threads) and
// then resets the count and bumps the generation the others wait for (one
// barrier per level instead of one before and one after the bookkeeping).
// The fences make the writes of every group before the barrier, and of
// `body`, visible to every group after it.
template <class Body>
__device__ __forceinline__ void grid_barrier_last(LevelCtl *ctl, uint32_t n_groups, Body body) {
  __shared__ uint32_t s_last, s_gen;
  __syncthreads();
  if (threadIdx.x == 0) {
    s_gen = __hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    s_last = atomicAdd(&ctl->bar_count, 1u) == n_groups - 1;
  }
  __syncthreads();
  if (s_last) {  // (uniform over the work-group)
    __threadfence();
    body();
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      __hip_atomic_fetch_add(&ctl->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    // relaxed polls (no cache invalidation per poll), one acquire after
    while (__hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == s_gen)
      __builtin_amdgcn_s_sleep(2);
    __threadfence();
  }
  __syncthreads();
}

template <typename PtrT, int WG>
__global__ __launch_bounds__(WG) void k_coop_levels(KArgs base, LevelCtl *ctl, CoopArgs c) {
  __shared__ LlbShared<PtrT, WG> sh;
  uint32_t cur = c.cur;
  for (uint32_t it = 0; it < c.max_levels; it++) {
    // the same control block for every work-group (written before the
    // launch or before the last barrier): the decision to stay is uniform
    const uint32_t F = ctl->q_count[cur];
    const unsigned long long EF = ctl->q_edges[cur];
    const int choice = c.fixed >= 0 ? c.fixed : (int)ctl->choice;
    if (F == 0 || c.mark.marks(EF, F) || (choice != BFSX_LLB && choice != BFSX_SLB)) break;
    KArgs a = base;
    a.q_in = c.queue[cur];
    a.q_out = c.queue[cur ^ 1];
    a.F = F;
    a.q_out_count = &ctl->q_count[cur ^ 1];
    a.q_out_edges = &ctl->q_edges[cur ^ 1];
    a.next_level = (int32_t)ctl->level + 1;
    a.mark = nullptr;
    const uint32_t n_tiles = (F + WG - 1) / WG;
    unsigned long long dacc = 0;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
      llb_tile<PtrT, WG>(a, sh, tile, n_tiles, choice == BFSX_SLB, dacc);
    block_atomic_add<WG>(dacc, a.q_out_edges);
    grid_barrier_last(ctl, gridDim.x, [&] {
      level_end_body(ctl, cur, cur ^ 1, 0, choice, c.Wg, c.n_cols, c.nnz, c.n_iso, c.allow, 1,
                     SpecGuard{nullptr, 0, 0, NEVER_MARKS});
    });
    cur ^= 1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) write_level_end(ctl, cur, c.mirror);
}

// sum of out-degree over reached vertices < n_rows, reached count, max level
template <typename PtrT>
__global__ void k_summarize(const PtrT *col_ptr, const int32_t *levels, uint32_t n_rows,
                            uint32_t n_cols, unsigned long long *out) {
  unsigned long long edges = 0, reached = 0;
  int32_t mx = -1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cols;
       i += gridDim.x * blockDim.x) {
    const int32_t l = levels[i];
    if (l >= 0) {
      reached++;
      mx = l > mx ? l : mx;
      if (i < n_rows) edges += (unsigned long long)(col_ptr[i + 1] - col_ptr[i]);
    }
  }
  edges = wave_sum(edges);
  reached = wave_sum(reached);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const int32_t o = __shfl_xor(mx, d, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    if (edges) atomicAdd(&out[0], edges);
    if (reached) atomicAdd(&out[1], reached);
    if (mx >= 0) atomicMax(&out[2], (unsigned long long)mx);
  }
}

// per-level vertex counts and out-degree sums; LDS histogram per block when
// the level count is small, global atomics otherwise
constexpr uint32_t HIST_LDS_LEVELS = 1024;

template <typename PtrT>
__global__ __launch_bounds__(256) void k_level_hist(const PtrT *col_ptr, const int32_t *levels,
                                                    uint32_t n_rows, uint32_t n_cols,
                                                    uint32_t max_levels,
                                                    unsigned long long *counts,
                                                    unsigned long long *edges) {
  __shared__ unsigned long long s_cnt[HIST_LDS_LEVELS];
  __shared__ unsigned long long s_edg[HIST_LDS_LEVELS];
  const bool lds = max_levels <= HIST_LDS_LEVELS;
  if (lds)
    for (uint32_t i = threadIdx.x; i < max_levels; i += blockDim.x) s_cnt[i] = s_edg[i] = 0;
  __syncthreads();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n_cols;
       v += gridDim.x * blockDim.x) {
    const int32_t l = levels[v];
    if (l < 0 || (uint32_t)l >= max_levels) continue;
    const unsigned long long d =
        v < n_rows ? (unsigned long long)(col_ptr[v + 1] - col_ptr[v]) : 0ull;
    if (lds) {
      atomicAdd(&s_cnt[l], 1ull);
      if (d) atomicAdd(&s_edg[l], d);
    } else {
      atomicAdd(&counts[l], 1ull);
      if (d) atomicAdd(&edges[l], d);
    }
  }
  __syncthreads();
  if (lds)
    for (uint32_t i = threadIdx.x; i < max_levels; i += blockDim.x) {
      if (s_cnt[i]) atomicAdd(&counts[i], s_cnt[i]);
      if (s_edg[i]) atomicAdd(&edges[i], s_edg[i]);
    }
}

__global__ void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                              uint64_t n16) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// Event pair around a step's main kernel (timed runs only).
struct KernelTimer {
  const StepArgs &s;
  hipStream_t st;
  KernelTimer(const StepArgs &a, hipStream_t stream) : s(a), st(stream) {
    if (s.ev_k0) BFSX_HIP(hipEventRecord(s.ev_k0, st));
  }
  ~KernelTimer() {
    if (s.ev_k1) (void)hipEventRecord(s.ev_k1, st);
  }
};

static unsigned cap_grid(uint64_t blocks, unsigned cap) {
  if (blocks == 0) return 1;
  return blocks > cap ? cap : (unsigned)blocks;
}

static KArgs kargs(const StepArgs &s, uint32_t hub) {
  KArgs a;
  a.col_ptr = s.col_ptr;
  a.row_idx = s.row_idx;
  a.n_rows = s.n_rows;
  a.n_cols = s.n_cols;
  a.q_in = s.q_in;
  a.F = s.F;
  a.q_out = s.q_out;
  a.q_out_count = &s.ctl->q_count[s.q_out_idx];
  a.q_out_edges = s.track_degrees ? &s.ctl->q_edges[s.q_out_idx] : nullptr;
  a.vis = s.vis;
  a.levels = s.levels;
  a.parents = s.parents;
  a.next_level = s.next_level;
  a.hub_degree = hub;
  a.mark = s.mark;
  a.examined = s.count_examined ? &s.ctl->examined : nullptr;
  a.hot_vis = nullptr;
  a.hot_id = nullptr;
  return a;
}

// hot_vis bit h = visited bit of hot_id[h] (H a multiple of 64, and so is
// the grid-stride: every wave runs whole rounds).
__global__ void k_hot_bits(const uint32_t *vis, const uint32_t *hot_id, uint32_t H,
                           uint32_t *hot_vis) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < H; h += stride) {
    const uint32_t v = hot_id[h];
    const unsigned long long m = __ballot((vis[v >> 5] >> (v & 31)) & 1u);
    if (lane == 0) hot_vis[h >> 5] = (uint32_t)m;
    if (lane == 32) hot_vis[h >> 5] = (uint32_t)(m >> 32);
  }
}

template <typename PtrT, bool STRIDED, bool HOT>
static void launch_llb_t(bfsx_ctx *ctx, const KArgs &a) {
  const uint32_t wg = ctx->params.llb_wg;
  const unsigned grid = (unsigned)((a.F + wg - 1) / wg);
  hipStream_t s = ctx->stream;
  switch (wg) {
    case 64: hipLaunchKernelGGL((k_llb<PtrT, 64, STRIDED, HOT>), dim3(grid), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((k_llb<PtrT, 128, STRIDED, HOT>), dim3(grid), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((k_llb<PtrT, 256, STRIDED, HOT>), dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((k_llb<PtrT, 512, STRIDED, HOT>), dim3(grid), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((k_llb<PtrT, 1024, STRIDED, HOT>), dim3(grid), dim3(1024), 0, s, a); break;
    default: fail(BFSX_ERR_ARG, "llb_wg must be 64, 128, 256, 512 or 1024");
  }
}

void step_llb(bfsx_ctx *ctx, const StepArgs &s, bool strided) {
  if (s.F == 0) return;
  KArgs a = kargs(s, ctx->params.hub_degree);
  if (a.mark && s.hot_ok && ctx->hot_n && s.row_idx == ctx->row_idx.as<uint32_t>()) {
    // a marking level probes the hot vertices' visited bits in an
    // L2-resident bitmap, gathered first.  (Before most of them are visited,
    // a hot probe that finds its vertex unvisited costs a dependent load of
    // the vertex id: hot_ok is false then.)
    a.row_idx = ctx->row_idx_hot.as<uint32_t>();
    a.hot_vis = ctx->hot_vis.as<uint32_t>();
    a.hot_id = ctx->hot_id.as<uint32_t>();
    hipLaunchKernelGGL(k_hot_bits, dim3(cap_grid(ctx->hot_n / 256 + 1, (unsigned)ctx->cu_count * 8)),
                       dim3(256), 0, ctx->stream, s.vis, a.hot_id, ctx->hot_n,
                       ctx->hot_vis.as<uint32_t>());
  }
  KernelTimer kt(s, ctx->stream);
  const bool hot = a.hot_vis != nullptr;
  if (ctx->ptr_bytes == 8) {
    if (hot) {
      if (strided) launch_llb_t<uint64_t, true, true>(ctx, a);
      else launch_llb_t<uint64_t, false, true>(ctx, a);
    } else {
      if (strided) launch_llb_t<uint64_t, true, false>(ctx, a);
      else launch_llb_t<uint64_t, false, false>(ctx, a);
    }
  } else {
    if (hot) {
      if (strided) launch_llb_t<uint32_t, true, true>(ctx, a);
      else launch_llb_t<uint32_t, false, true>(ctx, a);
    } else {
      if (strided) launch_llb_t<uint32_t, true, false>(ctx, a);
      else launch_llb_t<uint32_t, false, false>(ctx, a);
    }
  }
  BFSX_HIP(hipGetLastError());
}

// Bucketed discovery of a heavy GLB level (see k_glb_bucket).
template <typename PtrT>
static void glb_bucketed(bfsx_ctx *ctx, const StepArgs &s, const KArgs &a, const PtrT *g_scan,
                         const PtrT *g_start, uint32_t chunk, unsigned grid, uint32_t n_slices) {
  hipStream_t st = ctx->stream;
  // the frontier's edge count (kept exactly by every step) bounds the targets
  const uint64_t cap = s.EF ? s.EF : 1;
  ensure(ctx->bkt, sizeof(uint16_t) * cap, st);
  ensure(ctx->bkt_cnt, sizeof(uint32_t) * (size_t)grid * n_slices, st);
  ensure(ctx->bkt_tot, sizeof(uint64_t) * n_slices, st);
  ensure(ctx->bkt_base, sizeof(uint64_t) * (n_slices + 1), st);
  uint32_t *cnt = ctx->bkt_cnt.as<uint32_t>();
  uint64_t *tot = ctx->bkt_tot.as<uint64_t>();
  uint64_t *base = ctx->bkt_base.as<uint64_t>();
  uint16_t *bkt = ctx->bkt.as<uint16_t>();
  const uint32_t *clo = ctx->glb_chunk_lo.as<uint32_t>();
  KernelTimer kt(s, st);
  hipLaunchKernelGGL((k_glb_bucket<PtrT, BKT_WG, 1>), dim3(grid), dim3(BKT_WG), 0, st, a, g_scan,
                     g_start, clo, s.ctl, chunk, n_slices, cnt, (const uint64_t *)nullptr,
                     (uint16_t *)nullptr, (uint64_t)0);
  if (n_slices >= 256)
    hipLaunchKernelGGL((k_bucket_colscan<1024>), dim3((n_slices + 15) / 16), dim3(1024), 0, st,
                       cnt, (uint32_t)grid, n_slices, tot);
  else
    hipLaunchKernelGGL((k_bucket_colscan_few<256>), dim3(n_slices), dim3(256), 0, st, cnt,
                       (uint32_t)grid, n_slices, tot);
  hipLaunchKernelGGL(k_bucket_base, dim3(1), dim3(1024), 0, st, tot, n_slices, base);
  hipLaunchKernelGGL((k_glb_bucket<PtrT, BKT_WG, 2>), dim3(grid), dim3(BKT_WG), 0, st, a, g_scan,
                     g_start, clo, s.ctl, chunk, n_slices, cnt, base, bkt, cap);
  hipLaunchKernelGGL((k_bucket_apply<PtrT, 512>), dim3(n_slices), dim3(512), 0, st,
                     static_cast<const PtrT *>(s.col_ptr), ctx->deg16.as<uint16_t>(), s.n_rows,
                     s.n_cols, bkt, base, cap, s.vis,
                     s.levels, s.next_level, s.q_out, &s.ctl->q_count[s.q_out_idx],
                     s.track_degrees ? &s.ctl->q_edges[s.q_out_idx] : nullptr, s.fbm_next);
  BFSX_HIP(hipGetLastError());
}

template <typename PtrT>
static bool step_glb_t(bfsx_ctx *ctx, const StepArgs &s) {
  hipStream_t st = ctx->stream;
  const uint32_t F = s.F;
  const unsigned grid = (unsigned)ctx->cu_count * ctx->params.glb_blocks_per_cu;
  const uint32_t n_slices = (uint32_t)(((uint64_t)ctx->n_cols + (1u << BKT_SHIFT) - 1) >> BKT_SHIFT);
  const bool bucketed = s.bucket_ok && !s.parents && n_slices <= BKT_MAX;
  // Edges per work-queue item: glb_chunk, or, when the frontier's edges
  // (exact, kept on the device) would not give every expand work-group an
  // item, as few as GLB_MIN_CHUNK so that they spread over the grid (a
  // frontier of two hubs otherwise runs on a handful of work-groups).
  uint32_t chunk = bucketed ? BKT_CHUNK : ctx->params.glb_chunk;
  if (!bucketed && s.EF > 0 && s.EF < (unsigned long long)chunk * grid) {
    // (a power of two: the chunk table's divisions become shifts)
    const unsigned long long want = (s.EF + grid - 1) / grid;
    uint32_t c = GLB_MIN_CHUNK;
    while (c < want && c < chunk) c <<= 1;
    chunk = std::min(c, chunk);
  }
  const uint32_t n_tiles = (F + GLB_TILE - 1) / GLB_TILE;
  PtrT *g_loc = ctx->glb_scan.as<PtrT>();             // F entries, then F+1 global
  PtrT *g_scan = g_loc + ((uint64_t)ctx->n_cols + 1);
  PtrT *g_start = ctx->glb_start.as<PtrT>();
  PtrT *tile = ctx->glb_tile.as<PtrT>();
  PtrT *g_end = tile + ((uint64_t)ctx->n_cols / GLB_TILE + 2);
  uint32_t shift = 0;  // log2(chunk) when chunk is a power of two
  if ((chunk & (chunk - 1)) == 0)
    while ((1u << shift) < chunk) shift++;
  if (F <= GLB_SMALL_F) {
    hipLaunchKernelGGL(k_glb_scan_small<PtrT>, dim3(1), dim3(GLB_SMALL_WG), 0, st,
                       static_cast<const PtrT *>(s.col_ptr), s.n_rows, s.q_in, F, g_start, g_scan,
                       chunk, shift, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl);
  } else {
    hipLaunchKernelGGL(k_glb_tiles<PtrT>, dim3(n_tiles), dim3(256), 0, st,
                       static_cast<const PtrT *>(s.col_ptr), s.n_rows, s.q_in, F, g_loc, g_start,
                       tile);
    hipLaunchKernelGGL(k_glb_tile_offsets<PtrT>, dim3(1), dim3(1024), 0, st, tile, n_tiles,
                       chunk, s.ctl, g_end);
    hipLaunchKernelGGL(k_glb_offsets<PtrT>, dim3(cap_grid((F + 255) / 256, 65535)), dim3(256),
                       0, st, g_loc, tile, g_end, F, g_scan, chunk, shift,
                       ctx->glb_chunk_lo.as<uint32_t>());
  }
  const KArgs a = kargs(s, 0);
  if (bucketed) {
    // fewer, larger groups than the GLB expand grid: longer runs per
    // (group, slice) in the segments that pass 2 writes
    glb_bucketed<PtrT>(ctx, s, a, g_scan, g_start, chunk, (unsigned)ctx->cu_count * 4, n_slices);
    return true;
  }
  KernelTimer kt(s, st);
  const bool stage = !a.mark && chunk <= GLB_STAGE_CAP;
  switch (ctx->params.glb_wg) {
    case 128:
      if (stage)
        hipLaunchKernelGGL((k_glb_expand<PtrT, 128, true>), dim3(grid), dim3(128), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      else
        hipLaunchKernelGGL((k_glb_expand<PtrT, 128, false>), dim3(grid), dim3(128), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      break;
    case 256:
      if (stage)
        hipLaunchKernelGGL((k_glb_expand<PtrT, 256, true>), dim3(grid), dim3(256), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      else
        hipLaunchKernelGGL((k_glb_expand<PtrT, 256, false>), dim3(grid), dim3(256), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      break;
    case 512:
      if (stage)
        hipLaunchKernelGGL((k_glb_expand<PtrT, 512, true>), dim3(grid), dim3(512), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      else
        hipLaunchKernelGGL((k_glb_expand<PtrT, 512, false>), dim3(grid), dim3(512), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      break;
    case 1024:
      if (stage)
        hipLaunchKernelGGL((k_glb_expand<PtrT, 1024, true>), dim3(grid), dim3(1024), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      else
        hipLaunchKernelGGL((k_glb_expand<PtrT, 1024, false>), dim3(grid), dim3(1024), 0, st, a,
                           g_scan, g_start, ctx->glb_chunk_lo.as<uint32_t>(), s.ctl, chunk);
      break;
    default: fail(BFSX_ERR_ARG, "glb_wg must be 128, 256, 512 or 1024");
  }
  return false;
}

bool step_glb(bfsx_ctx *ctx, const StepArgs &s) {
  if (s.F == 0) return false;
  const uint64_t pw = ctx->ptr_bytes;
  // scratch: local + global scans, starts, tile sums, chunk table
  ensure(ctx->glb_scan, pw * 2 * ((uint64_t)ctx->n_cols + 1), ctx->stream);
  ensure(ctx->glb_start, pw * ((uint64_t)ctx->n_cols + 1), ctx->stream);
  ensure(ctx->glb_tile, pw * ((uint64_t)ctx->n_cols / GLB_TILE + 4), ctx->stream);
  ensure(ctx->glb_chunk_lo, sizeof(uint32_t) * (ctx->nnz / GLB_MIN_CHUNK + 2), ctx->stream);
  const bool bucketed =
      ctx->ptr_bytes == 8 ? step_glb_t<uint64_t>(ctx, s) : step_glb_t<uint32_t>(ctx, s);
  BFSX_HIP(hipGetLastError());
  return bucketed;
}

// Summary resolution for n vertices: the smallest shift >= 5 with at most
// max_words * 32 summary bits.
static uint32_t pull_summary_shift(uint64_t n, uint32_t max_words) {
  uint32_t shift = 5;
  while (((n + (1ull << shift) - 1) >> shift) > (uint64_t)max_words * 32) shift++;
  return shift;
}

// The two-phase kernel wherever the first-neighbour array exists.
template <typename PtrT, int WG>
static void launch_pull(unsigned grid, size_t lds, hipStream_t st, const KArgs &a,
                        const StepArgs &s, const uint32_t *g_sum, uint32_t sum_words,
                        uint32_t shift, const uint32_t *fd, const uint16_t *fdd,
                        const uint32_t *iso,
                        const uint32_t *crank, const uint32_t *hbm, const uint32_t *hub,
                        SpecGuard guard) {
  if (fd) {
    if (g_sum)
      hipLaunchKernelGGL((k_pull2<PtrT, WG, true>), dim3(grid), dim3(WG), lds, st, a, s.fbm_cur,
                         s.fbm_next, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard);
    else
      hipLaunchKernelGGL((k_pull2<PtrT, WG, false>), dim3(grid), dim3(WG), lds, st, a, s.fbm_cur,
                         s.fbm_next, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard);
  } else {
    if (g_sum)
      hipLaunchKernelGGL((k_pull<PtrT, WG, true>), dim3(grid), dim3(WG), lds, st, a, s.fbm_cur,
                         s.fbm_next, g_sum, sum_words, shift, guard);
    else
      hipLaunchKernelGGL((k_pull<PtrT, WG, false>), dim3(grid), dim3(WG), lds, st, a, s.fbm_cur,
                         s.fbm_next, g_sum, sum_words, shift, guard);
  }
}

template <typename PtrT>
static void step_pull_t(bfsx_ctx *ctx, const StepArgs &s, bool speculative) {
  const KArgs a = kargs(s, 0);
  const uint32_t wg = ctx->params.pull_wg;
  const uint32_t n_chunks = (s.n_rows + PULL_CHUNK - 1) / PULL_CHUNK;
  hipStream_t st = ctx->stream;
  // The summary fills the LDS that the kernel's own arrays leave free (at
  // most PULL_SUM_WORDS_MAX words), and pays when at most a quarter of its
  // bits can be set.
  // (static LDS: the found and undecided words of every wave, and the rest)
  const uint32_t static_lds = 2 * (wg / 64) * 64 * 4 + 1024;
  uint32_t max_words = ctx->lds_per_block > static_lds ? (ctx->lds_per_block - static_lds) / 4 : 0;
  max_words = max_words > PULL_SUM_WORDS_MAX ? PULL_SUM_WORDS_MAX : max_words;
  const uint32_t shift = pull_summary_shift(s.n_cols, max_words ? max_words : 1);
  const uint32_t sum_bits = (uint32_t)(((uint64_t)s.n_cols + (1ull << shift) - 1) >> shift);
  const uint32_t sum_words = (sum_bits + 31) / 32;
  // (a speculative step does not know its frontier's size: no summary)
  const bool use_sum = !speculative && max_words >= 64 &&
                       ((uint64_t)s.F << shift) * 4 <= (uint64_t)s.n_cols;
  const SpecGuard guard{speculative ? s.ctl : nullptr, s.q_in_idx, 1u << BFSX_PULL, NEVER_MARKS};
  const uint32_t *g_sum = nullptr;
  if (use_sum) {
    ensure(ctx->fsum, sizeof(uint32_t) * PULL_SUM_WORDS_MAX, st);
    const uint32_t n_words = (uint32_t)bitmap_words(s.n_cols);
    const unsigned sgrid = cap_grid(((uint64_t)sum_bits + 255) / 256, (unsigned)ctx->cu_count * 4);
    hipLaunchKernelGGL(k_fbm_summary, dim3(sgrid), dim3(256), 0, st, s.fbm_cur, n_words, shift,
                       ctx->fsum.as<uint32_t>(), sum_bits);
    g_sum = ctx->fsum.as<uint32_t>();
  }
  const size_t lds = use_sum ? sizeof(uint32_t) * sum_words : 0;
  // resident work-groups per CU: at most 32 waves, and what LDS allows
  const uint32_t nw = wg / 64;
  uint32_t per_cu = 32 / nw;
  const uint32_t by_lds = ctx->lds_per_cu / (uint32_t)(static_lds + lds);
  per_cu = by_lds < per_cu ? (by_lds ? by_lds : 1) : per_cu;
  const unsigned grid = cap_grid(((uint64_t)n_chunks + nw - 1) / nw,
                                 (unsigned)ctx->cu_count * per_cu);
  // the {first neighbour, degree} array belongs to the hub-first copy
  const uint32_t *fd = ctx->pull_fd.p && s.row_idx == ctx->row_idx_pull.as<uint32_t>()
                           ? ctx->pull_fd.as<uint32_t>()
                           : nullptr;
  const uint16_t *fdd = fd ? ctx->pull_fdd.as<uint16_t>() : nullptr;
  // (its entries exist for the vertices with neighbours: the others must
  // start visited, which a search does on a symmetric graph)
  if (fd && !((ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols))
    fail(BFSX_ERR_UNSUPPORTED, "bottom-up step on a non-symmetric graph");
  const uint32_t *iso = fd ? ctx->iso_bm.as<uint32_t>() : nullptr;
  const uint32_t *crank = fd ? ctx->pull_crank.as<uint32_t>() : nullptr;
  // hub indexing belongs to fd: the hubs' frontier bits, gathered first
  const uint32_t *hbm = nullptr, *hub = nullptr;
  if (fd && ctx->pull_hubs) {
    hub = ctx->pull_hub.as<uint32_t>();
    hbm = ctx->pull_hbm.as<uint32_t>();
    hipLaunchKernelGGL(k_hub_bits, dim3(cap_grid(ctx->pull_hubs / 256 + 1, (unsigned)ctx->cu_count * 4)),
                       dim3(256), 0, st, s.fbm_cur, hub, ctx->pull_hubs, ctx->pull_hbm.as<uint32_t>(),
                       guard);
  }
  KernelTimer kt(s, st);
  switch (wg) {
    case 64: launch_pull<PtrT, 64>(grid, lds, st, a, s, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard); break;
    case 128: launch_pull<PtrT, 128>(grid, lds, st, a, s, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard); break;
    case 256: launch_pull<PtrT, 256>(grid, lds, st, a, s, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard); break;
    case 512: launch_pull<PtrT, 512>(grid, lds, st, a, s, g_sum, sum_words, shift, fd, fdd, iso, crank, hbm, hub, guard); break;
    default: fail(BFSX_ERR_ARG, "pull_wg must be 64, 128, 256 or 512");
  }
}

void step_pull(bfsx_ctx *ctx, const StepArgs &s, bool speculative) {
  if (ctx->ptr_bytes == 8) step_pull_t<uint64_t>(ctx, s, speculative);
  else step_pull_t<uint32_t>(ctx, s, speculative);
  BFSX_HIP(hipGetLastError());
}

// Compaction launch shape: about eight groups per CU, each owning between
// 4 (one chunk per wave) and COMPACT_MAX_CHUNKS consecutive chunks.
struct CompactShape {
  uint32_t n_chunks, per_group;
  unsigned grid;
};
static CompactShape compact_shape(const bfsx_ctx *ctx, uint32_t n_cols) {
  CompactShape c;
  c.n_chunks = (uint32_t)(((uint64_t)n_cols + 1023) / 1024);
  const uint64_t groups = (uint64_t)ctx->cu_count * 8;
  uint64_t per = (c.n_chunks + groups - 1) / groups;
  per = per < COMPACT_WG / 64 ? COMPACT_WG / 64 : per;
  per = per > COMPACT_MAX_CHUNKS ? COMPACT_MAX_CHUNKS : per;
  c.per_group = (uint32_t)per;
  c.grid = (unsigned)((c.n_chunks + per - 1) / per);
  if (c.grid == 0) c.grid = 1;
  return c;
}

void compact(bfsx_ctx *ctx, const StepArgs &s, uint8_t *mark, uint32_t *fbm_next) {
  const CompactShape c = compact_shape(ctx, s.n_cols);
  unsigned long long *qe = s.track_degrees ? &s.ctl->q_edges[s.q_out_idx] : nullptr;
  uint32_t *qc = &s.ctl->q_count[s.q_out_idx];
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_compact<uint64_t>, dim3(c.grid), dim3(COMPACT_WG), 0, ctx->stream,
                       static_cast<const uint64_t *>(s.col_ptr), ctx->deg16.as<uint16_t>(),
                       s.n_rows, mark, c.n_chunks,
                       c.per_group, s.vis, s.levels, s.next_level, s.q_out, qc, qe, fbm_next);
  else
    hipLaunchKernelGGL(k_compact<uint32_t>, dim3(c.grid), dim3(COMPACT_WG), 0, ctx->stream,
                       static_cast<const uint32_t *>(s.col_ptr), ctx->deg16.as<uint16_t>(),
                       s.n_rows, mark, c.n_chunks,
                       c.per_group, s.vis, s.levels, s.next_level, s.q_out, qc, qe, fbm_next);
  BFSX_HIP(hipGetLastError());
}

// Frontier queue `q` of the bitmap `bm` (after a bottom-up step); the slot
// cursor is the control block's conversion counter, zeroed by level_end.
void bitmap_to_queue(bfsx_ctx *ctx, const uint32_t *bm, uint32_t *q) {
  const uint32_t n_words = (uint32_t)bitmap_words(ctx->n_cols);
  const uint64_t groups = (uint64_t)ctx->cu_count * 8;
  uint32_t per = (uint32_t)((n_words + groups - 1) / groups);
  per = (per + 255) & ~255u;  // whole 1 KiB runs of words per group
  const unsigned grid = (unsigned)((n_words + per - 1) / per);
  LevelCtl *ctl = ctx->ctl.as<LevelCtl>();
  hipLaunchKernelGGL((k_bm_to_queue<256>), dim3(grid), dim3(256), 0, ctx->stream, bm, n_words,
                     per, q, &ctl->conv);
  BFSX_HIP(hipGetLastError());
}

void queue_to_fbm(bfsx_ctx *ctx, const uint32_t *q, uint32_t F, uint32_t *fbm) {
  const uint32_t n_words = (uint32_t)bitmap_words(ctx->n_cols);
  hipLaunchKernelGGL(k_fill_words, dim3(cap_grid(((uint64_t)n_words + 255) / 256, (unsigned)ctx->cu_count * 8)),
                     dim3(256), 0, ctx->stream, fbm, n_words);
  if (F)
    hipLaunchKernelGGL(k_queue_bits, dim3(cap_grid(((uint64_t)F + 255) / 256, (unsigned)ctx->cu_count * 8)),
                       dim3(256), 0, ctx->stream, q, F, fbm);
  BFSX_HIP(hipGetLastError());
}


void init_search(bfsx_ctx *ctx, uint32_t root, bool symmetric_iso, uint32_t *par) {
  const uint32_t n_words = (uint32_t)bitmap_words(ctx->n_cols);
  const unsigned grid = cap_grid(((uint64_t)ctx->n_cols + 255) / 256,
                                 (unsigned)ctx->cu_count * 16);
  const uint32_t *iso = symmetric_iso ? ctx->iso_bm.as<uint32_t>() : nullptr;
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_init<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint64_t>(), ctx->n_rows, ctx->n_cols, root,
                       ctx->levels.as<int32_t>(), par, ctx->vis.as<uint32_t>(), iso, n_words,
                       ctx->queue[0].as<uint32_t>(), ctx->ctl.as<LevelCtl>());
  else
    hipLaunchKernelGGL(k_init<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint32_t>(), ctx->n_rows, ctx->n_cols, root,
                       ctx->levels.as<int32_t>(), par, ctx->vis.as<uint32_t>(), iso, n_words,
                       ctx->queue[0].as<uint32_t>(), ctx->ctl.as<LevelCtl>());
  BFSX_HIP(hipGetLastError());
}

void level_begin(bfsx_ctx *ctx) {
  hipLaunchKernelGGL(k_level_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->ctl.as<LevelCtl>());
  BFSX_HIP(hipGetLastError());
}

void level_end(bfsx_ctx *ctx, uint32_t consumed, uint32_t produced, bool stamp, int kind,
               uint32_t decide_allow, bool reuse, uint32_t spec_kinds,
               const MarkRule *spec_heavy) {
  const uint64_t n_iso = (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols
                             ? ctx->n_isolated
                             : 0;
  hipLaunchKernelGGL(k_level_end, dim3(1), dim3(64), 0, ctx->stream, ctx->ctl.as<LevelCtl>(),
                     consumed, produced, stamp ? 1 : 0, kind,
                     decide_allow ? ctx->sched_w.as<float>() : (const float *)nullptr,
                     (uint64_t)ctx->n_cols, (uint64_t)ctx->nnz, n_iso, decide_allow,
                     reuse ? 1 : 0, ctx->d_mirror,
                     SpecGuard{spec_kinds ? ctx->ctl.as<LevelCtl>() : nullptr, consumed,
                               spec_kinds, spec_heavy ? *spec_heavy : NEVER_MARKS});
  BFSX_HIP(hipGetLastError());
}

// One device-driven top-down level (LLB or SLB, or the choice in the control
// block when fixed < 0) for a captured loop: step, compaction, bookkeeping.
template <typename PtrT>
static void dev_level_t(bfsx_ctx *ctx, uint32_t cur, int fixed, bool track, MarkRule mark,
                        bool speculative) {
  KArgs a{};
  a.col_ptr = ctx->col_ptr.p;
  a.row_idx = ctx->row_idx.as<uint32_t>();
  a.n_rows = ctx->n_rows;
  a.n_cols = ctx->n_cols;
  a.q_in = ctx->queue[cur].as<uint32_t>();
  a.q_out = ctx->queue[cur ^ 1].as<uint32_t>();
  LevelCtl *ctl = ctx->ctl.as<LevelCtl>();
  a.q_out_count = &ctl->q_count[cur ^ 1];
  a.q_out_edges = track ? &ctl->q_edges[cur ^ 1] : nullptr;
  a.vis = ctx->vis.as<uint32_t>();
  a.levels = ctx->levels.as<int32_t>();
  a.parents = nullptr;
  a.hub_degree = ctx->params.hub_degree;
  a.mark = ctx->mark.as<uint8_t>();
  const DevStep d{cur, fixed, mark, speculative ? 1 : 0};
  const uint32_t wg = ctx->params.llb_wg;
  const unsigned grid = cap_grid(((uint64_t)ctx->n_cols + wg - 1) / wg, (unsigned)ctx->cu_count * 8);
  hipStream_t s = ctx->stream;
  switch (wg) {
    case 64: hipLaunchKernelGGL((k_td_dev<PtrT, 64>), dim3(grid), dim3(64), 0, s, a, ctl, d); break;
    case 128: hipLaunchKernelGGL((k_td_dev<PtrT, 128>), dim3(grid), dim3(128), 0, s, a, ctl, d); break;
    case 256: hipLaunchKernelGGL((k_td_dev<PtrT, 256>), dim3(grid), dim3(256), 0, s, a, ctl, d); break;
    case 512: hipLaunchKernelGGL((k_td_dev<PtrT, 512>), dim3(grid), dim3(512), 0, s, a, ctl, d); break;
    case 1024: hipLaunchKernelGGL((k_td_dev<PtrT, 1024>), dim3(grid), dim3(1024), 0, s, a, ctl, d); break;
    default: fail(BFSX_ERR_ARG, "llb_wg must be 64, 128, 256, 512 or 1024");
  }
  if (speculative) {  // (it never marks)
    BFSX_HIP(hipGetLastError());
    return;
  }
  const CompactShape c = compact_shape(ctx, ctx->n_cols);
  hipLaunchKernelGGL(k_compact_dev<PtrT>, dim3(c.grid), dim3(COMPACT_WG), 0, s,
                     static_cast<const PtrT *>(ctx->col_ptr.p), ctx->deg16.as<uint16_t>(),
                     ctx->n_rows,
                     ctx->mark.as<uint8_t>(), c.n_chunks, c.per_group, ctx->vis.as<uint32_t>(),
                     ctx->levels.as<int32_t>(), ctl, d, ctx->queue[cur ^ 1].as<uint32_t>(),
                     track ? 1 : 0, (uint32_t *)nullptr);
  BFSX_HIP(hipGetLastError());
}

void dev_level(bfsx_ctx *ctx, uint32_t cur, int fixed, bool track, MarkRule mark,
               bool speculative) {
  if (ctx->ptr_bytes == 8) dev_level_t<uint64_t>(ctx, cur, fixed, track, mark, speculative);
  else dev_level_t<uint32_t>(ctx, cur, fixed, track, mark, speculative);
}

template <typename PtrT, int WG>
static bool coop_launch(bfsx_ctx *ctx, const KArgs &a, const CoopArgs &c) {
  const void *f = reinterpret_cast<const void *>(&k_coop_levels<PtrT, WG>);
  int per_cu = 0;
  BFSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, WG, 0));
  if (per_cu < 1) return false;
  const unsigned groups = std::min<unsigned>(ctx->params.coop_groups, (unsigned)(ctx->cu_count * per_cu));
  const dim3 grid(groups), block(WG);
  KArgs ka = a;
  LevelCtl *ctl = ctx->ctl.as<LevelCtl>();
  CoopArgs ca = c;
  void *args[] = {&ka, &ctl, &ca};
  BFSX_HIP(hipLaunchCooperativeKernel(f, grid, block, args, 0, ctx->stream));
  return true;
}

template <typename PtrT>
static bool coop_levels_t(bfsx_ctx *ctx, const KArgs &a, const CoopArgs &c) {
  switch (ctx->params.llb_wg) {
    case 64: return coop_launch<PtrT, 64>(ctx, a, c);
    case 128: return coop_launch<PtrT, 128>(ctx, a, c);
    case 256: return coop_launch<PtrT, 256>(ctx, a, c);
    case 512: return coop_launch<PtrT, 512>(ctx, a, c);
    case 1024: return coop_launch<PtrT, 1024>(ctx, a, c);
    default: fail(BFSX_ERR_ARG, "llb_wg must be 64, 128, 256, 512 or 1024");
  }
}

bool coop_levels(bfsx_ctx *ctx, uint32_t cur, int fixed, uint32_t allow, bool learned,
                 MarkRule mark, uint32_t max_levels) {
  if (!ctx->coop_ok) return false;
  KArgs a{};
  a.col_ptr = ctx->col_ptr.p;
  a.row_idx = ctx->row_idx.as<uint32_t>();
  a.n_rows = ctx->n_rows;
  a.n_cols = ctx->n_cols;
  a.vis = ctx->vis.as<uint32_t>();
  a.levels = ctx->levels.as<int32_t>();
  a.parents = nullptr;
  a.hub_degree = ctx->params.hub_degree;
  a.mark = nullptr;
  a.examined = nullptr;
  CoopArgs c{};
  c.queue[0] = ctx->queue[0].as<uint32_t>();
  c.queue[1] = ctx->queue[1].as<uint32_t>();
  c.cur = cur;
  c.fixed = fixed;
  c.mark = mark;
  c.max_levels = max_levels;
  c.Wg = learned ? ctx->sched_w.as<float>() : nullptr;
  c.n_cols = ctx->n_cols;
  c.nnz = ctx->nnz;
  c.n_iso = (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols ? ctx->n_isolated : 0;
  c.allow = allow;
  c.mirror = ctx->d_mirror;
  return ctx->ptr_bytes == 8 ? coop_levels_t<uint64_t>(ctx, a, c) : coop_levels_t<uint32_t>(ctx, a, c);
}

void reached_summary(bfsx_ctx *ctx, unsigned long long *d_out) {
  const unsigned grid = cap_grid(((uint64_t)ctx->n_cols + 255) / 256,
                                 (unsigned)ctx->cu_count * 8);
  BFSX_HIP(hipMemsetAsync(d_out, 0, 3 * sizeof(unsigned long long), ctx->stream));
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_summarize<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint64_t>(), ctx->levels.as<int32_t>(), ctx->n_rows,
                       ctx->n_cols, d_out);
  else
    hipLaunchKernelGGL(k_summarize<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint32_t>(), ctx->levels.as<int32_t>(), ctx->n_rows,
                       ctx->n_cols, d_out);
  BFSX_HIP(hipGetLastError());
}

void level_histogram(bfsx_ctx *ctx, uint32_t max_levels, unsigned long long *d_counts,
                     unsigned long long *d_edges) {
  const unsigned grid = cap_grid(((uint64_t)ctx->n_cols + 255) / 256,
                                 (unsigned)ctx->cu_count * 4);
  BFSX_HIP(hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * max_levels, ctx->stream));
  BFSX_HIP(hipMemsetAsync(d_edges, 0, sizeof(unsigned long long) * max_levels, ctx->stream));
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_level_hist<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint64_t>(), ctx->levels.as<int32_t>(), ctx->n_rows,
                       ctx->n_cols, max_levels, d_counts, d_edges);
  else
    hipLaunchKernelGGL(k_level_hist<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint32_t>(), ctx->levels.as<int32_t>(), ctx->n_rows,
                       ctx->n_cols, max_levels, d_counts, d_edges);
  BFSX_HIP(hipGetLastError());
}

void stream_copy(bfsx_ctx *ctx, const void *src, void *dst, uint64_t bytes) {
  const uint64_t n16 = bytes / 16;
  const unsigned grid = cap_grid((n16 + 255) / 256, (unsigned)ctx->cu_count * 16);
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream,
                     static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
  BFSX_HIP(hipGetLastError());
}

}  // namespace bfsx
