// Synthetic graphs built directly in HBM.
//
// The reference makes its R-MAT edge lists with networkit on the host
// (scripts/genGraph.py:41-70) and converts them to CSC with scipy
// (scripts/generator.py:151-195).  Here the same kind of graph is generated
// on the GPU so that a scale-26 graph (about 2^31 stored edges) never crosses
// PCIe: a counter-based RNG gives every edge its own random stream, the
// permutation is a sort of random keys, and deduplication/symmetrization is a
// radix sort of packed (source, destination) keys followed by unique.  The
// CSC layout is the reference's: col_ptr over source vertices, row_idx holds
// each source's destinations in ascending order (scripts/generator.py:115-147
// writes scipy's CSC arrays, which are sorted).
//
// Random stream layout (ctr[3] is a stream tag; key = (lo(seed), hi(seed))):
//   tag 0  R-MAT edge k, quadrant choice j: ctr = (lo k, hi k, j/4, 0), word j%4
//   tag 1  permutation key of vertex x:    ctr = (x, 0, 0, 1)
//   tag 2  root candidate k:               ctr = (lo k, hi k, 0, 2)
//   tag 3  grid edge g keep/drop:          ctr = (lo g, hi g, 0, 3)
//   tag 4  grid long-range edge k:         ctr = (k, 0, 0, 4)
#include <cstring>
#include <utility>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "bfsx_internal.hpp"

namespace bfsx {

// Philox-4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64
// multiplies with the Weyl key schedule.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r) {
      k.x += 0x9E3779B9u;
      k.y += 0xBB67AE85u;
    }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
  }
  return c;
}

__device__ __forceinline__ uint2 seed_key(uint64_t seed) {
  return make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
}

__global__ void k_rmat_pairs(int scale, uint64_t m, uint32_t tA, uint32_t tAB,
                             uint32_t tABC, uint64_t seed, uint32_t *u,
                             uint32_t *v) {
  const uint2 key = seed_key(seed);
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = 0, b = 0;
    uint4 r = make_uint4(0, 0, 0, 0);
    for (int j = 0; j < scale; j++) {
      if ((j & 3) == 0)
        r = philox4x32_10(make_uint4((uint32_t)k, (uint32_t)(k >> 32), j >> 2, 0), key);
      const uint32_t x = (j & 3) == 0 ? r.x : (j & 3) == 1 ? r.y : (j & 3) == 2 ? r.z : r.w;
      const uint32_t bit = 1u << (scale - 1 - j);
      if (x >= tA) {
        if (x < tAB) {
          b |= bit;
        } else if (x < tABC) {
          a |= bit;
        } else {
          a |= bit;
          b |= bit;
        }
      }
    }
    u[k] = a;
    v[k] = b;
  }
}

__global__ void k_perm_keys(uint32_t n, int bits, uint64_t seed, uint64_t *keys) {
  const uint2 key = seed_key(seed);
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += gridDim.x * blockDim.x) {
    const uint4 r = philox4x32_10(make_uint4(x, 0, 0, 1), key);
    const uint64_t k64 = ((uint64_t)r.x << 32) | r.y;
    keys[x] = ((k64 >> bits) << bits) | x;
  }
}

__global__ void k_perm_scatter(uint32_t n, uint64_t mask, const uint64_t *sorted,
                               uint32_t *perm) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    perm[sorted[i] & mask] = i;
}

__global__ void k_apply_perm(uint64_t m, const uint32_t *perm, uint32_t offset, uint32_t *u,
                             uint32_t *v) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x) {
    u[k] = perm[u[k]] + offset;
    v[k] = perm[v[k]] + offset;
  }
}

__global__ void k_grid_pairs(uint32_t side, uint32_t t_drop, uint64_t seed,
                             uint32_t *u, uint32_t *v, uint32_t *count) {
  const uint2 key = seed_key(seed);
  const uint64_t H = (uint64_t)side * (side - 1);
  const uint64_t G = 2 * H;
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < G;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 r = philox4x32_10(make_uint4((uint32_t)g, (uint32_t)(g >> 32), 0, 3), key);
    uint64_t a, b;
    if (g < H) {
      a = (g / (side - 1)) * side + g % (side - 1);
      b = a + 1;
    } else {
      const uint64_t h = g - H;
      a = (h / side) * side + h % side;
      b = a + side;
    }
    const bool keep = r.x >= t_drop;
    // Dropped edges become self-loops, which the CSC builder removes; this
    // keeps the pair array dense without a compaction pass.
    u[g] = (uint32_t)a;
    v[g] = keep ? (uint32_t)b : (uint32_t)a;
    (void)count;
  }
}

__global__ void k_long_pairs(uint32_t n_long, uint64_t V, uint64_t seed,
                             uint32_t *u, uint32_t *v) {
  const uint2 key = seed_key(seed);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n_long;
       k += gridDim.x * blockDim.x) {
    const uint4 r = philox4x32_10(make_uint4(k, 0, 0, 4), key);
    u[k] = (uint32_t)(((uint64_t)r.x * V) >> 32);
    v[k] = (uint32_t)(((uint64_t)r.y * V) >> 32);
  }
}

// keys[2i], keys[2i+1] = both directions of pair i packed as src<<B | dst;
// self loops become the sentinel n<<B, which sorts after every real key.
__global__ void k_sym_keys(uint64_t m, int B, uint64_t sentinel, const uint32_t *u,
                           const uint32_t *v, uint64_t *keys) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t a = u[i], b = v[i];
    const bool loop = a == b;
    keys[2 * i] = loop ? sentinel : ((a << B) | b);
    keys[2 * i + 1] = loop ? sentinel : ((b << B) | a);
  }
}

// Splits unique sorted keys into row_idx and col_ptr.  col_ptr[x] = first key
// index with source >= x: entry i writes col_ptr for the sources in
// (src(i-1), src(i)]; the tail (last source, n] gets nnz.
template <typename PtrT>
__global__ void k_csc_from_keys(const uint64_t *uk, uint64_t nnz, int B, uint32_t n,
                                PtrT *col_ptr, uint32_t *row_idx) {
  const uint64_t mask = (B >= 64) ? ~0ull : ((1ull << B) - 1);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= nnz;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = (i < nnz) ? (uk[i] >> B) : (uint64_t)n + 1;
    const int64_t sp = (i == 0) ? -1 : (int64_t)(uk[i - 1] >> B);
    const uint64_t hi = (i < nnz) ? s : (uint64_t)n;
    for (uint64_t x = (uint64_t)(sp + 1); x <= hi; x++) col_ptr[x] = (PtrT)i;
    if (i < nnz) row_idx[i] = (uint32_t)(uk[i] & mask);
  }
}

static int bits_for(uint64_t n) {
  int b = 0;
  while ((1ull << b) < n) b++;
  return b;
}

static unsigned grid_for(uint64_t n, unsigned block = 256) {
  uint64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return (unsigned)g;
}

void gen_rmat_pairs(bfsx_ctx *ctx, int scale, uint64_t m, uint32_t tA, uint32_t tAB,
                    uint32_t tABC, uint64_t seed_graph, uint64_t seed_perm,
                    uint32_t *d_u, uint32_t *d_v) {
  gen_rmat_pairs_offset(ctx, scale, m, tA, tAB, tABC, seed_graph, seed_perm, 0, d_u, d_v);
}

// Same pairs with every vertex id shifted by `offset` (graphs packed into
// one vertex range).
void gen_rmat_pairs_offset(bfsx_ctx *ctx, int scale, uint64_t m, uint32_t tA, uint32_t tAB,
                           uint32_t tABC, uint64_t seed_graph, uint64_t seed_perm,
                           uint32_t offset, uint32_t *d_u, uint32_t *d_v) {
  hipStream_t s = ctx->stream;
  const uint32_t n = 1u << scale;
  hipLaunchKernelGGL(k_rmat_pairs, dim3(grid_for(m)), dim3(256), 0, s, scale, m, tA,
                     tAB, tABC, seed_graph, d_u, d_v);
  BFSX_HIP(hipGetLastError());
  // permutation: perm[x] = rank of (random high bits | x)
  DevBuf keys, sorted, perm, tmp;
  ensure(keys, sizeof(uint64_t) * n, s);
  ensure(sorted, sizeof(uint64_t) * n, s);
  ensure(perm, sizeof(uint32_t) * n, s);
  const int bits = bits_for(n);
  hipLaunchKernelGGL(k_perm_keys, dim3(grid_for(n)), dim3(256), 0, s, n, bits, seed_perm,
                     keys.as<uint64_t>());
  size_t tbytes = 0;
  BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(),
                                    sorted.as<uint64_t>(), (size_t)n, 0, 64, s));
  ensure(tmp, tbytes, s);
  BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(),
                                    sorted.as<uint64_t>(), (size_t)n, 0, 64, s));
  const uint64_t mask = (1ull << bits) - 1;
  hipLaunchKernelGGL(k_perm_scatter, dim3(grid_for(n)), dim3(256), 0, s, n, mask,
                     sorted.as<uint64_t>(), perm.as<uint32_t>());
  hipLaunchKernelGGL(k_apply_perm, dim3(grid_for(m)), dim3(256), 0, s, m,
                     perm.as<uint32_t>(), offset, d_u, d_v);
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  release(keys);
  release(sorted);
  release(perm);
  release(tmp);
}

uint64_t gen_grid_pairs(bfsx_ctx *ctx, uint32_t side, uint32_t t_drop, uint32_t n_long,
                        uint64_t seed, uint32_t *d_u, uint32_t *d_v) {
  hipStream_t s = ctx->stream;
  const uint64_t G = 2ull * side * (side - 1);
  const uint64_t V = (uint64_t)side * side;
  if (G)
    hipLaunchKernelGGL(k_grid_pairs, dim3(grid_for(G)), dim3(256), 0, s, side, t_drop,
                       seed, d_u, d_v, (uint32_t *)nullptr);
  if (n_long)
    hipLaunchKernelGGL(k_long_pairs, dim3(grid_for(n_long)), dim3(256), 0, s, n_long, V,
                       seed, d_u + G, d_v + G);
  BFSX_HIP(hipGetLastError());
  return G + n_long;
}

void build_symmetric_csc(bfsx_ctx *ctx, uint32_t n, uint32_t *d_u, uint32_t *d_v,
                         uint64_t m, int ptr_bytes) {
  hipStream_t s = ctx->stream;
  const int B = bits_for(n) > 0 ? bits_for(n) : 1;
  const uint64_t sentinel = (uint64_t)n << B;
  const int end_bit = B + bits_for((uint64_t)n + 1);
  const uint64_t K = 2 * m;
  DevBuf keys, sorted, tmp, cnt;
  ensure(keys, sizeof(uint64_t) * (K ? K : 1), s);
  hipLaunchKernelGGL(k_sym_keys, dim3(grid_for(m)), dim3(256), 0, s, m, B, sentinel, d_u,
                     d_v, keys.as<uint64_t>());
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  // the pair arrays are no longer needed; the caller frees them after return
  ensure(sorted, sizeof(uint64_t) * (K ? K : 1), s);
  size_t tbytes = 0;
  BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(),
                                    sorted.as<uint64_t>(), (size_t)K, 0, end_bit, s));
  ensure(tmp, tbytes, s);
  BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(),
                                    sorted.as<uint64_t>(), (size_t)K, 0, end_bit, s));
  ensure(cnt, sizeof(unsigned long long), s);
  size_t ubytes = 0;
  BFSX_HIP(rocprim::unique(nullptr, ubytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                           cnt.as<unsigned long long>(), (size_t)K,
                           rocprim::equal_to<uint64_t>(), s));
  if (ubytes > tmp.bytes) {
    release(tmp);
    ensure(tmp, ubytes, s);
  }
  BFSX_HIP(rocprim::unique(tmp.p, ubytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                           cnt.as<unsigned long long>(), (size_t)K,
                           rocprim::equal_to<uint64_t>(), s));
  unsigned long long n_unique = 0;
  BFSX_HIP(hipMemcpyAsync(&n_unique, cnt.p, sizeof(n_unique), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  uint64_t nnz = n_unique;
  if (nnz) {
    uint64_t last = 0;
    BFSX_HIP(hipMemcpyAsync(&last, keys.as<uint64_t>() + nnz - 1, sizeof(last),
                            hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    if (last == sentinel) nnz--;
  }
  release(sorted);
  release(tmp);
  release(cnt);
  if (ptr_bytes == 4 && nnz > 0xFFFFFFFFull)
    fail(BFSX_ERR_OVERFLOW, "graph has more than 2^32-1 edges; use 8-byte col_ptr");
  release(ctx->col_ptr);
  release(ctx->row_idx);
  ensure(ctx->col_ptr, (size_t)ptr_bytes * ((uint64_t)n + 1), s);
  ensure(ctx->row_idx, sizeof(uint32_t) * (nnz ? nnz : 1), s);
  if (ptr_bytes == 8)
    hipLaunchKernelGGL(k_csc_from_keys<uint64_t>, dim3(grid_for(nnz + 1)), dim3(256), 0, s,
                       keys.as<uint64_t>(), nnz, B, n, ctx->col_ptr.as<uint64_t>(),
                       ctx->row_idx.as<uint32_t>());
  else
    hipLaunchKernelGGL(k_csc_from_keys<uint32_t>, dim3(grid_for(nnz + 1)), dim3(256), 0, s,
                       keys.as<uint64_t>(), nnz, B, n, ctx->col_ptr.as<uint32_t>(),
                       ctx->row_idx.as<uint32_t>());
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  release(keys);
  ctx->n_rows = ctx->n_cols = n;
  ctx->nnz = nnz;
  ctx->ptr_bytes = ptr_bytes;
  ctx->gflags = BFSX_GRAPH_SYMMETRIC;
  ctx->has_graph = true;
}

// ---- graph summary: max degree, isolated-vertex bitmap and count ---------
template <typename PtrT>
__global__ void k_summary(const PtrT *col_ptr, uint32_t n_rows, uint32_t n_cols,
                          uint32_t *iso_bm, unsigned int *max_deg,
                          unsigned int *n_iso) {
  const uint32_t nwords = (uint32_t)bitmap_words(n_cols);
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += gridDim.x * blockDim.x) {
    uint32_t bits = 0, mx = 0;
    for (int b = 0; b < 32; b++) {
      const uint32_t x = w * 32 + b;
      if (x < n_rows) {
        const uint64_t d = (uint64_t)(col_ptr[x + 1] - col_ptr[x]);
        const uint32_t dc = d > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)d;
        if (dc == 0) bits |= 1u << b;
        mx = dc > mx ? dc : mx;
      }
    }
    iso_bm[w] = bits;
    if (bits) atomicAdd(n_iso, (unsigned)__popc(bits));
    if (mx) atomicMax(max_deg, mx);
  }
}

template <typename PtrT>
__global__ void k_deg16(const PtrT *col_ptr, uint32_t n_rows, uint64_t n_pad, uint16_t *deg16) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n_pad;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = v < n_rows ? (uint64_t)(col_ptr[v + 1] - col_ptr[v]) : 0ull;
    deg16[v] = (uint16_t)(d < DEG16_ESCAPE ? d : DEG16_ESCAPE);
  }
}

void graph_summary(bfsx_ctx *ctx) {
  hipStream_t s = ctx->stream;
  const uint32_t nwords = (uint32_t)bitmap_words(ctx->n_cols);
  const uint64_t n_pad = (uint64_t)(nwords ? nwords : 1) * 32;
  ensure(ctx->deg16, sizeof(uint16_t) * n_pad, s);
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_deg16<uint64_t>, dim3(grid_for(n_pad)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint64_t>(), ctx->n_rows, n_pad, ctx->deg16.as<uint16_t>());
  else
    hipLaunchKernelGGL(k_deg16<uint32_t>, dim3(grid_for(n_pad)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint32_t>(), ctx->n_rows, n_pad, ctx->deg16.as<uint16_t>());
  ensure(ctx->iso_bm, sizeof(uint32_t) * (nwords ? nwords : 2), s);
  DevBuf out;
  ensure(out, 2 * sizeof(unsigned int), s);
  BFSX_HIP(hipMemsetAsync(out.p, 0, 2 * sizeof(unsigned int), s));
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_summary<uint64_t>, dim3(grid_for(nwords)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint64_t>(), ctx->n_rows, ctx->n_cols,
                       ctx->iso_bm.as<uint32_t>(), out.as<unsigned int>(),
                       out.as<unsigned int>() + 1);
  else
    hipLaunchKernelGGL(k_summary<uint32_t>, dim3(grid_for(nwords)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint32_t>(), ctx->n_rows, ctx->n_cols,
                       ctx->iso_bm.as<uint32_t>(), out.as<unsigned int>(),
                       out.as<unsigned int>() + 1);
  BFSX_HIP(hipGetLastError());
  unsigned int h[2] = {0, 0};
  BFSX_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  release(out);
  ctx->max_degree = h[0];
  ctx->n_isolated = h[1];
}

// ---- adjacency order --------------------------------------------------------
// Each column's list is re-sorted in place by a vertex rank: rank(v) = v
// (ascending ids, the builder's order) or the position of v when vertices
// are sorted by (degree descending, id ascending) ("hub first").  A
// bottom-up step stops at the first neighbour in the frontier, and on
// power-law graphs the frontier holds the hubs early on, so hub-first lists
// end most scans after one probe.  Levels do not depend on the order.
template <typename PtrT>
__global__ void k_degree_keys(const PtrT *col_ptr, uint32_t n_rows, uint32_t n_cols,
                              uint32_t max_deg, int B, uint64_t *keys) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n_cols;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = v < n_rows ? (uint64_t)(col_ptr[v + 1] - col_ptr[v]) : 0ull;
    const uint64_t dc = d > max_deg ? max_deg : d;
    keys[v] = ((uint64_t)(max_deg - dc) << B) | v;
  }
}

// order[i] = vertex of rank i, rank[v] = rank of v
__global__ void k_ranks(const uint64_t *sorted, uint32_t n, uint64_t mask, uint32_t *order,
                        uint32_t *rank) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(sorted[i] & mask);
    order[i] = v;
    rank[v] = (uint32_t)i;
  }
}

// key of edge e in column c: (c << B) | rank(row_idx[e]); a wave per column
template <typename PtrT>
__global__ void k_edge_rank_keys(const PtrT *col_ptr, uint32_t n_rows, const uint32_t *row_idx,
                                 const uint32_t *rank, int B, uint64_t *keys) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t c = blockIdx.x * (uint64_t)(blockDim.x / 64) + (threadIdx.x >> 6); c < n_rows;
       c += nwaves) {
    const PtrT b = col_ptr[c], e = col_ptr[c + 1];
    for (PtrT k = b + lane; k < e; k += 64)
      keys[k] = (c << B) | (rank ? rank[row_idx[k]] : row_idx[k]);
  }
}

__global__ void k_rows_from_keys(const uint64_t *sorted, uint64_t nnz, uint64_t mask,
                                 const uint32_t *order, uint32_t *row_idx) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nnz;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(sorted[k] & mask);
    row_idx[k] = order ? order[r] : r;
  }
}

template <typename PtrT>
static void order_adjacency_t(bfsx_ctx *ctx, bool hub_first, uint32_t *out) {
  hipStream_t s = ctx->stream;
  const uint32_t n = ctx->n_cols;
  const uint64_t nnz = ctx->nnz;
  const int B = bits_for(n) > 0 ? bits_for(n) : 1;
  const uint64_t mask = (1ull << B) - 1;
  const PtrT *cp = ctx->col_ptr.as<PtrT>();
  DevBuf keys, sorted, tmp, rank, order;
  struct Releaser {  // temporaries go on every exit path
    DevBuf *b[5];
    ~Releaser() {
      for (DevBuf *x : b) release(*x);
    }
  } releaser{{&keys, &sorted, &tmp, &rank, &order}};
  size_t tbytes = 0;
  if (hub_first) {
    ensure(keys, sizeof(uint64_t) * n, s);
    ensure(sorted, sizeof(uint64_t) * n, s);
    ensure(rank, sizeof(uint32_t) * n, s);
    ensure(order, sizeof(uint32_t) * n, s);
    hipLaunchKernelGGL(k_degree_keys<PtrT>, dim3(grid_for(n)), dim3(256), 0, s, cp,
                       ctx->n_rows, n, ctx->max_degree, B, keys.as<uint64_t>());
    BFSX_HIP(hipGetLastError());
    const int end_bit = B + bits_for((uint64_t)ctx->max_degree + 1);
    BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(),
                                      sorted.as<uint64_t>(), (size_t)n, 0, end_bit, s));
    ensure(tmp, tbytes, s);
    BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(),
                                      sorted.as<uint64_t>(), (size_t)n, 0, end_bit, s));
    hipLaunchKernelGGL(k_ranks, dim3(grid_for(n)), dim3(256), 0, s, sorted.as<uint64_t>(), n,
                       mask, order.as<uint32_t>(), rank.as<uint32_t>());
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipStreamSynchronize(s));
    release(keys);
    release(sorted);
    release(tmp);
  }
  if (nnz) {
    ensure(keys, sizeof(uint64_t) * nnz, s);
    hipLaunchKernelGGL(k_edge_rank_keys<PtrT>, dim3(grid_for((uint64_t)ctx->n_rows * 64)),
                       dim3(256), 0, s, cp, ctx->n_rows, ctx->row_idx.as<uint32_t>(),
                       hub_first ? rank.as<uint32_t>() : (const uint32_t *)nullptr, B,
                       keys.as<uint64_t>());
    BFSX_HIP(hipGetLastError());
    ensure(sorted, sizeof(uint64_t) * nnz, s);
    const int end_bit = B + bits_for((uint64_t)ctx->n_rows + 1);
    tbytes = 0;
    BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(),
                                      sorted.as<uint64_t>(), (size_t)nnz, 0, end_bit, s));
    ensure(tmp, tbytes, s);
    BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(),
                                      sorted.as<uint64_t>(), (size_t)nnz, 0, end_bit, s));
    hipLaunchKernelGGL(k_rows_from_keys, dim3(grid_for(nnz)), dim3(256), 0, s,
                       sorted.as<uint64_t>(), nnz, mask,
                       hub_first ? order.as<uint32_t>() : (const uint32_t *)nullptr, out);
    BFSX_HIP(hipGetLastError());
  }
  BFSX_HIP(hipStreamSynchronize(s));
}

template <typename PtrT>
__global__ void k_first_deg(const PtrT *col_ptr, const uint32_t *row_idx, uint32_t n_rows,
                            uint2 *fd) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n_rows;
       v += gridDim.x * blockDim.x) {
    const PtrT b = col_ptr[v], e = col_ptr[v + 1];
    fd[v] = make_uint2(e > b ? row_idx[b] : ~0u, (uint32_t)(e - b));
  }
}

void first_neighbours(bfsx_ctx *ctx, const uint32_t *row_idx, uint2 *fd) {
  const uint32_t n = ctx->n_rows;
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_first_deg<uint64_t>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint64_t>(), row_idx, n, fd);
  else
    hipLaunchKernelGGL(k_first_deg<uint32_t>, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                       ctx->col_ptr.as<uint32_t>(), row_idx, n, fd);
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(ctx->stream));
}

// hrank[v] = h for the hub hub[h] (hrank was filled with ~0u)
__global__ void k_hub_rank(const uint32_t *hub, uint32_t H, uint32_t *hrank) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < H; h += gridDim.x * blockDim.x)
    hrank[hub[h]] = h;
}

// first neighbours that are hubs: their index, flagged
__global__ void k_fd_hub_index(uint2 *fd, uint32_t n_rows, const uint32_t *hrank) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n_rows;
       v += gridDim.x * blockDim.x) {
    const uint2 x = fd[v];
    if (x.y == 0) continue;
    const uint32_t h = hrank[x.x];
    if (h != ~0u) fd[v].x = PULL_HUB_FLAG | h;
  }
}

void index_hub_neighbours(bfsx_ctx *ctx) {
  release(ctx->pull_hub);
  release(ctx->pull_hbm);
  ctx->pull_hubs = 0;
  const uint32_t n = ctx->n_rows;
  if (!ctx->pull_fd.p || ctx->n_cols >= PULL_HUB_FLAG || n < 64) return;
  // about one vertex in 256, in whole waves of 64
  uint64_t H = ((uint64_t)n / 256 + 63) & ~63ull;
  H = std::max<uint64_t>(H, 64);
  H = std::min<uint64_t>(H, PULL_HUBS_MAX);  // (and H <= n: n >= 64)
  hipStream_t s = ctx->stream;
  ensure(ctx->pull_hub, sizeof(uint32_t) * H, s);
  ensure(ctx->pull_hbm, sizeof(uint32_t) * (H / 32), s);
  top_degree_vertices(ctx, (uint32_t)H, ctx->pull_hub.as<uint32_t>());
  DevBuf hrank;
  ensure(hrank, sizeof(uint32_t) * ctx->n_cols, s);
  BFSX_HIP(hipMemsetAsync(hrank.p, 0xFF, sizeof(uint32_t) * ctx->n_cols, s));
  hipLaunchKernelGGL(k_hub_rank, dim3(grid_for(H)), dim3(256), 0, s,
                     ctx->pull_hub.as<uint32_t>(), (uint32_t)H, hrank.as<uint32_t>());
  hipLaunchKernelGGL(k_fd_hub_index, dim3(grid_for(n)), dim3(256), 0, s,
                     ctx->pull_fd.as<uint2>(), n, hrank.as<uint32_t>());
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  release(hrank);
  ctx->pull_hubs = (uint32_t)H;
}

// Vertices with neighbours (below n_rows) in word w of the isolated bitmap.
__device__ __forceinline__ uint32_t linked_word(const uint32_t *iso, uint32_t w, uint32_t n_rows) {
  const uint64_t v0 = (uint64_t)w * 32;
  if (v0 >= n_rows) return 0u;
  const uint32_t m = n_rows - v0 < 32 ? (1u << (n_rows - v0)) - 1u : ~0u;
  return ~iso[w] & m;
}

// cnt[c] = vertices with neighbours in chunk c (a thread per chunk)
__global__ void k_chunk_linked(const uint32_t *iso, uint32_t n_rows, uint32_t n_chunks,
                               uint32_t *cnt) {
  constexpr uint32_t WPC = PULL_CHUNK / 32;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += gridDim.x * blockDim.x) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < WPC; i++) n += __popc(linked_word(iso, c * WPC + i, n_rows));
    cnt[c] = n;
  }
}

// In-place exclusive scan of cnt[0..n) (one work-group of 1024 threads,
// each owning a contiguous run); the total lands in cnt[n].
__global__ __launch_bounds__(1024) void k_scan_runs(uint32_t *cnt, uint32_t n) {
  __shared__ uint64_t s_tot[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t run = (n + 1023) / 1024;
  const uint32_t i0 = min(t * run, n), i1 = min(i0 + run, n);
  uint64_t sum = 0;
  for (uint32_t i = i0; i < i1; i++) sum += cnt[i];
  s_tot[t] = sum;
  __syncthreads();
  if (t == 0) {  // (one-off, 1024 entries)
    uint64_t acc = 0;
    for (uint32_t k = 0; k < 1024; k++) {
      const uint64_t x = s_tot[k];
      s_tot[k] = acc;
      acc += x;
    }
    cnt[n] = (uint32_t)acc;
  }
  __syncthreads();
  uint64_t pos = s_tot[t];
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t x = cnt[i];
    cnt[i] = (uint32_t)pos;
    pos += x;
  }
}

// out[crank[c] + k] = fd[v].x and outd[...] = min(fd[v].y, DEG16_ESCAPE) for
// the k-th vertex v with neighbours of chunk c
__global__ void k_fd_compact(const uint2 *fd, const uint32_t *iso, uint32_t n_rows,
                             uint32_t n_chunks, const uint32_t *crank, uint32_t *out,
                             uint16_t *outd) {
  constexpr uint32_t WPC = PULL_CHUNK / 32;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += gridDim.x * blockDim.x) {
    uint32_t pos = crank[c];
    for (uint32_t i = 0; i < WPC; i++) {
      uint32_t m = linked_word(iso, c * WPC + i, n_rows);
      while (m) {
        const uint32_t b = __ffs(m) - 1;
        m &= m - 1;
        const uint2 x = fd[(c * WPC + i) * 32 + b];
        out[pos] = x.x;
        outd[pos] = (uint16_t)(x.y < DEG16_ESCAPE ? x.y : DEG16_ESCAPE);
        pos++;
      }
    }
  }
}

void compact_first_neighbours(bfsx_ctx *ctx) {
  release(ctx->pull_crank);
  release(ctx->pull_fdd);
  if (!ctx->pull_fd.p) return;
  hipStream_t s = ctx->stream;
  const uint32_t n = ctx->n_rows;
  const uint32_t n_chunks = (uint32_t)(((uint64_t)n + PULL_CHUNK - 1) / PULL_CHUNK);
  ensure(ctx->pull_crank, sizeof(uint32_t) * ((uint64_t)n_chunks + 1), s);
  uint32_t *crank = ctx->pull_crank.as<uint32_t>();
  const uint32_t *iso = ctx->iso_bm.as<uint32_t>();
  hipLaunchKernelGGL(k_chunk_linked, dim3(grid_for(n_chunks)), dim3(256), 0, s, iso, n,
                     n_chunks, crank);
  hipLaunchKernelGGL(k_scan_runs, dim3(1), dim3(1024), 0, s, crank, n_chunks);
  BFSX_HIP(hipGetLastError());
  uint32_t linked = 0;
  BFSX_HIP(hipMemcpyAsync(&linked, crank + n_chunks, sizeof(linked), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  if (linked != n - ctx->n_isolated) fail(BFSX_ERR_DEVICE, "isolated-vertex count mismatch");
  DevBuf out;
  ensure(out, sizeof(uint32_t) * ((uint64_t)linked + 1), s);
  ensure(ctx->pull_fdd, sizeof(uint16_t) * ((uint64_t)linked + 1), s);
  // the padding entry: a vertex with no neighbours (read, never used)
  BFSX_HIP(hipMemsetAsync(out.as<uint32_t>() + linked, 0, sizeof(uint32_t), s));
  BFSX_HIP(hipMemsetAsync(ctx->pull_fdd.as<uint16_t>() + linked, 0, sizeof(uint16_t), s));
  hipLaunchKernelGGL(k_fd_compact, dim3(grid_for(n_chunks)), dim3(256), 0, s,
                     ctx->pull_fd.as<uint2>(), iso, n, n_chunks, crank, out.as<uint32_t>(),
                     ctx->pull_fdd.as<uint16_t>());
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  std::swap(ctx->pull_fd, out);
  release(out);
}

// out[k] = PULL_HUB_FLAG | hrank[x] for hot x = row_idx[k], else x
__global__ void k_flag_hot(const uint32_t *row_idx, uint64_t nnz, const uint32_t *hrank,
                           uint32_t *out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride) {
    const uint32_t x = row_idx[k];
    const uint32_t h = hrank[x];
    out[k] = h != ~0u ? PULL_HUB_FLAG | h : x;
  }
}

void build_hot_copy(bfsx_ctx *ctx) {
  release(ctx->row_idx_hot);
  release(ctx->hot_id);
  release(ctx->hot_vis);
  ctx->hot_n = 0;
  const uint32_t n = ctx->n_rows;
  // Only where the visited bitmap outgrows an XCD's L2 (4 MiB) by enough
  // that most probes miss it: below 2^24 vertices (a 2 MiB bitmap) the
  // whole bitmap stays in L2 and the hot path only adds work (fixed LLB on
  // R-MAT 19 / 22 lost 22 % / 13 % with it).
  if (ctx->n_cols >= PULL_HUB_FLAG || ctx->n_cols < HOT_MIN_VERTICES || ctx->nnz == 0) return;
  // one vertex in 16, in whole waves of 64, at most 2^22 (a 512 KiB bitmap;
  // at R-MAT scale 26 one in 8 was no faster)
  uint64_t H = ((uint64_t)n / 16 + 63) & ~63ull;
  H = std::max<uint64_t>(H, 64);
  H = std::min<uint64_t>(H, 1ull << 22);  // (and H <= n)
  hipStream_t s = ctx->stream;
  ensure(ctx->hot_id, sizeof(uint32_t) * H, s);
  ensure(ctx->hot_vis, sizeof(uint32_t) * (H / 32), s);
  ensure(ctx->row_idx_hot, sizeof(uint32_t) * ctx->nnz, s);
  top_degree_vertices(ctx, (uint32_t)H, ctx->hot_id.as<uint32_t>());
  DevBuf hrank;
  ensure(hrank, sizeof(uint32_t) * ctx->n_cols, s);
  BFSX_HIP(hipMemsetAsync(hrank.p, 0xFF, sizeof(uint32_t) * ctx->n_cols, s));
  hipLaunchKernelGGL(k_hub_rank, dim3(grid_for(H)), dim3(256), 0, s, ctx->hot_id.as<uint32_t>(),
                     (uint32_t)H, hrank.as<uint32_t>());
  hipLaunchKernelGGL(k_flag_hot, dim3(grid_for(ctx->nnz)), dim3(256), 0, s,
                     ctx->row_idx.as<uint32_t>(), ctx->nnz, hrank.as<uint32_t>(),
                     ctx->row_idx_hot.as<uint32_t>());
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
  release(hrank);
  ctx->hot_n = (uint32_t)H;
}

void order_adjacency(bfsx_ctx *ctx, bool hub_first, uint32_t *out) {
  if (ctx->ptr_bytes == 8) order_adjacency_t<uint64_t>(ctx, hub_first, out);
  else order_adjacency_t<uint32_t>(ctx, hub_first, out);
}

// ---- seeded root candidates ------------------------------------------------
// Candidate k is x = (w0 * V) >> 32 of its tag-2 draw; ok[k] says whether it
// qualifies (degree >= 1, or reached in component_levels when given).
template <typename PtrT>
__global__ void k_root_candidates(const PtrT *col_ptr, uint32_t n_rows, uint32_t n_cols,
                                  uint64_t seed, uint64_t first, uint32_t count,
                                  const int32_t *comp, uint32_t *out_v, uint8_t *ok) {
  const uint2 key = seed_key(seed);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += gridDim.x * blockDim.x) {
    const uint64_t k = first + i;
    const uint4 r = philox4x32_10(make_uint4((uint32_t)k, (uint32_t)(k >> 32), 0, 2), key);
    const uint32_t x = (uint32_t)(((uint64_t)r.x * n_cols) >> 32);
    bool good;
    if (comp) good = comp[x] >= 0;
    else good = x < n_rows && col_ptr[x + 1] > col_ptr[x];
    out_v[i] = x;
    ok[i] = good ? 1 : 0;
  }
}

void root_candidates(bfsx_ctx *ctx, uint64_t seed, uint64_t first, uint32_t count,
                     const int32_t *d_comp, uint32_t *d_out_v, uint8_t *d_ok) {
  hipStream_t s = ctx->stream;
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_root_candidates<uint64_t>, dim3(grid_for(count)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint64_t>(), ctx->n_rows, ctx->n_cols, seed, first,
                       count, d_comp, d_out_v, d_ok);
  else
    hipLaunchKernelGGL(k_root_candidates<uint32_t>, dim3(grid_for(count)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint32_t>(), ctx->n_rows, ctx->n_cols, seed, first,
                       count, d_comp, d_out_v, d_ok);
  BFSX_HIP(hipGetLastError());
}

// ---- degree relabelling (batched searches) ----------------------------------
// new id i = rank of v when vertices are sorted by (degree descending, id
// ascending); the relabelled lists are ascending in new ids, i.e. hub first.
// On a power-law graph the hubs, which most lists reach, then sit in the
// lowest ids, so the per-vertex bitsets a bottom-up level gathers from them
// share a few L2-resident megabytes.
template <typename PtrT>
__global__ void k_relabel_degrees(const PtrT *col_ptr, const uint32_t *order, uint32_t n,
                                  PtrT *d) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x)
    d[i] = i < n ? col_ptr[order[i] + 1] - col_ptr[order[i]] : PtrT(0);
}

// key of edge k of old column c: (rank[c] << B) | rank[row_idx[k]]; a wave per column
template <typename PtrT>
__global__ void k_relabel_keys(const PtrT *col_ptr, uint32_t n, const uint32_t *row_idx,
                               const uint32_t *rank, int B, uint64_t *keys) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t c = blockIdx.x * (uint64_t)(blockDim.x / 64) + (threadIdx.x >> 6); c < n;
       c += nwaves) {
    const PtrT b = col_ptr[c], e = col_ptr[c + 1];
    const uint64_t hi = (uint64_t)rank[c] << B;
    for (PtrT k = b + lane; k < e; k += 64) keys[k] = hi | rank[row_idx[k]];
  }
}

__global__ void k_low_bits(const uint64_t *keys, uint64_t m, uint64_t mask, uint32_t *out) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x)
    out[k] = (uint32_t)(keys[k] & mask);
}

template <typename PtrT>
static void relabel_t(bfsx_ctx *ctx, uint32_t *rank, uint32_t *order, PtrT *new_cp,
                      uint32_t *new_ri) {
  hipStream_t s = ctx->stream;
  const uint32_t n = ctx->n_cols;
  const uint64_t nnz = ctx->nnz;
  const int B = bits_for(n) > 0 ? bits_for(n) : 1;
  const uint64_t mask = (1ull << B) - 1;
  const PtrT *cp = ctx->col_ptr.as<PtrT>();
  DevBuf keys, sorted, tmp, deg;
  struct Releaser {
    DevBuf *b[4];
    ~Releaser() {
      for (DevBuf *x : b) release(*x);
    }
  } releaser{{&keys, &sorted, &tmp, &deg}};
  // vertex order
  ensure(keys, sizeof(uint64_t) * n, s);
  ensure(sorted, sizeof(uint64_t) * n, s);
  hipLaunchKernelGGL(k_degree_keys<PtrT>, dim3(grid_for(n)), dim3(256), 0, s, cp, n, n,
                     ctx->max_degree, B, keys.as<uint64_t>());
  size_t tbytes = 0;
  const int vend = B + bits_for((uint64_t)ctx->max_degree + 1);
  BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    (size_t)n, 0, vend, s));
  ensure(tmp, tbytes, s);
  BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    (size_t)n, 0, vend, s));
  hipLaunchKernelGGL(k_ranks, dim3(grid_for(n)), dim3(256), 0, s, sorted.as<uint64_t>(), n, mask,
                     order, rank);
  // relabelled col_ptr: exclusive scan of the degrees in the new order
  ensure(deg, sizeof(PtrT) * ((uint64_t)n + 1), s);
  hipLaunchKernelGGL(k_relabel_degrees<PtrT>, dim3(grid_for((uint64_t)n + 1)), dim3(256), 0, s,
                     cp, order, n, deg.as<PtrT>());
  tbytes = 0;
  BFSX_HIP(rocprim::exclusive_scan(nullptr, tbytes, deg.as<PtrT>(), new_cp, PtrT(0),
                                   (size_t)n + 1, rocprim::plus<PtrT>(), s));
  ensure(tmp, tbytes, s);
  BFSX_HIP(rocprim::exclusive_scan(tmp.p, tbytes, deg.as<PtrT>(), new_cp, PtrT(0),
                                   (size_t)n + 1, rocprim::plus<PtrT>(), s));
  BFSX_HIP(hipStreamSynchronize(s));
  release(keys);
  release(sorted);
  // relabelled lists: sort the edges by (new column, new neighbour)
  if (nnz) {
    ensure(keys, sizeof(uint64_t) * nnz, s);
    ensure(sorted, sizeof(uint64_t) * nnz, s);
    hipLaunchKernelGGL(k_relabel_keys<PtrT>, dim3(grid_for((uint64_t)n * 64)), dim3(256), 0, s,
                       cp, n, ctx->row_idx.as<uint32_t>(), rank, B, keys.as<uint64_t>());
    tbytes = 0;
    BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(),
                                      sorted.as<uint64_t>(), (size_t)nnz, 0, 2 * B, s));
    ensure(tmp, tbytes, s);
    BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                      (size_t)nnz, 0, 2 * B, s));
    hipLaunchKernelGGL(k_low_bits, dim3(grid_for(nnz)), dim3(256), 0, s, sorted.as<uint64_t>(),
                       nnz, mask, new_ri);
  }
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
}

void relabel_by_degree(bfsx_ctx *ctx, uint32_t *rank, uint32_t *order, void *new_cp,
                       uint32_t *new_ri) {
  if (ctx->n_rows != ctx->n_cols) fail(BFSX_ERR_UNSUPPORTED, "relabelling needs n_rows == n_cols");
  if (ctx->ptr_bytes == 8)
    relabel_t<uint64_t>(ctx, rank, order, static_cast<uint64_t *>(new_cp), new_ri);
  else
    relabel_t<uint32_t>(ctx, rank, order, static_cast<uint32_t *>(new_cp), new_ri);
}

// Top-k by degree (scripts/rootNodes.py:133-137: sorted(G.degree, reverse)
// [:k], a stable sort, so equal degrees keep ascending ids): one 64-bit key
// per vertex < n_rows, (~degree << 32 | id) for degrees below 2^32; an
// ascending radix sort puts the largest degrees first and, among equal
// degrees, the smaller id first.
template <typename PtrT>
__global__ void k_degree_rank_keys(const PtrT *col_ptr, uint32_t n_rows, uint64_t *keys) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n_rows;
       v += gridDim.x * blockDim.x) {
    const uint64_t d = (uint64_t)(col_ptr[v + 1] - col_ptr[v]);
    const uint32_t dc = d > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)d;
    keys[v] = ((uint64_t)(~dc) << 32) | v;
  }
}

__global__ void k_low_words(const uint64_t *keys, uint32_t k, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = (uint32_t)keys[i];
}

void top_degree_vertices(bfsx_ctx *ctx, uint32_t k, uint32_t *d_out) {
  hipStream_t s = ctx->stream;
  const uint32_t n = ctx->n_rows;
  if (k > n) k = n;
  if (k == 0) return;
  DevBuf keys, sorted, tmp;
  ensure(keys, sizeof(uint64_t) * n, s);
  ensure(sorted, sizeof(uint64_t) * n, s);
  if (ctx->ptr_bytes == 8)
    hipLaunchKernelGGL(k_degree_rank_keys<uint64_t>, dim3(grid_for(n)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint64_t>(), n, keys.as<uint64_t>());
  else
    hipLaunchKernelGGL(k_degree_rank_keys<uint32_t>, dim3(grid_for(n)), dim3(256), 0, s,
                       ctx->col_ptr.as<uint32_t>(), n, keys.as<uint64_t>());
  size_t tbytes = 0;
  BFSX_HIP(rocprim::radix_sort_keys(nullptr, tbytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    (size_t)n, 0, 64, s));
  ensure(tmp, tbytes, s);
  BFSX_HIP(rocprim::radix_sort_keys(tmp.p, tbytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                    (size_t)n, 0, 64, s));
  hipLaunchKernelGGL(k_low_words, dim3(grid_for(k)), dim3(256), 0, s, sorted.as<uint64_t>(), k,
                     d_out);
  BFSX_HIP(hipGetLastError());
  BFSX_HIP(hipStreamSynchronize(s));
}

}  // namespace bfsx
