// Allocation-free single-source search: bfsx_run_noalloc.
//
// A variant of bfsx_run for callers that cannot allocate device memory.  The
// caller holds the CSC graph and the level array in device memory; the
// search state -- the visited bitmap, the two frontier queues and their
// counters -- lives in `scratch`, a fixed device array of this file (one copy
// per GPU, held until the process ends).  Host code never passes the array's
// name to a kernel: its device address is looked up once per device with
// hipGetSymbolAddress and kept in na_base[].  No call here allocates.
//
// Every level is one top-down expansion with exact-once discovery (an
// atomicOr on the visited bitmap elects the one lane that writes the level
// and appends the vertex; one atomicAdd per wave for the queue slots), a
// wave per frontier vertex, followed by a 4-byte copy of the next frontier's
// size to the host.  The state is re-initialised by every call (visited
// bitmap, counters, root), so calls do not depend on what an earlier call
// left in the array; calls that overlap on one GPU race on it.
#include <mutex>
#include <string>

#include "bfsx_internal.hpp"

namespace bfsx {
namespace {

constexpr uint32_t NA_MAX = BFSX_NOALLOC_MAX_VERTICES;
constexpr size_t NA_CTL_BYTES = 256;                          // counters
constexpr size_t NA_VIS_BYTES = (size_t)NA_MAX / 8;           // visited bitmap
constexpr size_t NA_Q_BYTES = ((size_t)(NA_MAX + 1) * 4 + 255) / 256 * 256;  // one queue
constexpr size_t NA_BYTES = NA_CTL_BYTES + NA_VIS_BYTES + 2 * NA_Q_BYTES;
static_assert(NA_BYTES == BFSX_NOALLOC_SCRATCH_BYTES, "scratch size in bfsx.h");

__device__ __align__(256) uint8_t scratch[NA_BYTES];

struct NaCtl {
  uint32_t count[2];  // entries in queue 0 / 1
  uint32_t bad;       // a col_ptr or row_idx entry out of range was seen
};

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
void *na_base[NA_MAX_DEVICES];  // device address of scratch, per device

__device__ __forceinline__ uint32_t na_lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// levels -1 except the root (0), visited = {root}, queue 0 = {root}
__global__ void k_na_init(int32_t *levels, uint32_t n_cols, uint32_t root, uint32_t *vis,
                          uint32_t n_words, uint32_t *q0, NaCtl *ctl) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cols; i += stride)
    levels[i] = i == root ? 0 : -1;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n_words; w += stride)
    vis[w] = w == (root >> 5) ? 1u << (root & 31) : 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q0[0] = root;
    ctl->count[0] = 1;
    ctl->count[1] = 0;
    ctl->bad = 0;
  }
}

__global__ void k_na_zero(uint32_t *p) {
  if (threadIdx.x == 0) *p = 0;
}

// One wave per frontier vertex (grid-stride over the F entries of q_in); its
// lanes take 64 consecutive edges per round.  Entries outside the graph
// (col_ptr beyond nnz or decreasing, row_idx >= n_cols) are skipped and
// flagged, never dereferenced.
template <typename PtrT>
__global__ __launch_bounds__(256) void k_na_expand(const PtrT *col_ptr, const uint32_t *row_idx,
                                                   uint64_t nnz, uint32_t n_rows, uint32_t n_cols,
                                                   const uint32_t *q_in, uint32_t F,
                                                   uint32_t *q_out, uint32_t *count_out,
                                                   uint32_t *vis, int32_t *levels,
                                                   int32_t next_level, uint32_t *bad) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < F; i += nwaves) {
    const uint32_t v = q_in[i];
    if (v >= n_rows) continue;  // (wave-uniform)
    uint64_t b = (uint64_t)col_ptr[v], e = (uint64_t)col_ptr[v + 1];
    if (e > nnz || b > e) {
      if (lane == 0) atomicOr(bad, 1u);
      e = e > nnz ? nnz : e;
      b = b > e ? e : b;
    }
    for (uint64_t base = b; base < e; base += 64) {
      const uint64_t k = base + lane;
      uint32_t n = k < e ? row_idx[k] : 0u;
      bool in = k < e;
      if (in && n >= n_cols) {
        atomicOr(bad, 1u);
        in = false;
      }
      bool isnew = false;
      if (in) {
        const uint32_t w = n >> 5, bit = 1u << (n & 31);
        isnew = !(vis[w] & bit) && !(atomicOr(&vis[w], bit) & bit);
      }
      if (isnew) levels[n] = next_level;
      const unsigned long long m = __ballot(isnew);
      if (m) {
        const int leader = __ffsll(m) - 1;
        uint32_t slot = 0;
        if ((int)lane == leader) slot = atomicAdd(count_out, (uint32_t)__popcll(m));
        slot = __shfl(slot, leader, 64);
        if (isnew) q_out[slot + na_lanes_below(m)] = n;
      }
    }
  }
}

// Device address of this GPU's scratch array (looked up once per device,
// with that device current).
void *na_scratch(int device) {
  std::lock_guard<std::mutex> g(na_mu);
  if (!na_base[device]) {
    void *p = nullptr;
    BFSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(BFSX_ERR_DEVICE, "scratch array has no device address");
    na_base[device] = p;
  }
  return na_base[device];
}

template <typename PtrT>
uint32_t na_search(uint8_t *base, hipStream_t s, const PtrT *col_ptr, const uint32_t *row_idx,
                   uint64_t nnz, uint32_t n_cols, uint32_t n_rows, uint32_t root,
                   int32_t *levels, int cus) {
  NaCtl *ctl = reinterpret_cast<NaCtl *>(base);
  uint32_t *vis = reinterpret_cast<uint32_t *>(base + NA_CTL_BYTES);
  uint32_t *q[2] = {reinterpret_cast<uint32_t *>(base + NA_CTL_BYTES + NA_VIS_BYTES),
                    reinterpret_cast<uint32_t *>(base + NA_CTL_BYTES + NA_VIS_BYTES + NA_Q_BYTES)};
  const uint32_t n_words = (n_cols + 31) / 32;
  const unsigned cap = (unsigned)cus * 8;
  auto grid = [&](uint64_t items, uint32_t per) {
    const uint64_t g = (items + per - 1) / per;
    return (unsigned)(g == 0 ? 1 : g > cap ? cap : g);
  };
  hipLaunchKernelGGL(k_na_init, dim3(grid(n_cols, 256)), dim3(256), 0, s, levels, n_cols, root,
                     vis, n_words, q[0], ctl);
  BFSX_HIP(hipGetLastError());
  uint32_t F = 1, cur = 0, level = 0;
  while (F > 0) {
    if (level > n_cols) fail(BFSX_ERR_DEVICE, "level loop did not terminate");
    if (level > 0) hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(64), 0, s, &ctl->count[cur ^ 1]);
    hipLaunchKernelGGL(k_na_expand<PtrT>, dim3(grid((uint64_t)F * 64, 256)), dim3(256), 0, s,
                       col_ptr, row_idx, nnz, n_rows, n_cols, q[cur], F, q[cur ^ 1],
                       &ctl->count[cur ^ 1], vis, levels, (int32_t)level + 1, &ctl->bad);
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(&F, &ctl->count[cur ^ 1], sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
    cur ^= 1;
    level++;
  }
  uint32_t bad = 0;
  BFSX_HIP(hipMemcpyAsync(&bad, &ctl->bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  BFSX_HIP(hipStreamSynchronize(s));
  if (bad) fail(BFSX_ERR_ARG, "col_ptr or row_idx entry out of range");
  return level;
}

thread_local std::string na_error;

}  // namespace

const char *noalloc_last_error() { return na_error.c_str(); }

}  // namespace bfsx

extern "C" int bfsx_run_noalloc(int device, void *stream, const void *col_ptr, int col_ptr_bytes,
                                const uint32_t *row_idx, uint32_t n_cols, uint32_t n_rows,
                                uint64_t nnz, uint32_t root, int32_t *levels,
                                uint32_t *n_levels) {
  using namespace bfsx;
  int old = -1;
  try {
    if (!col_ptr || !levels || (nnz && !row_idx)) fail(BFSX_ERR_ARG, "null device pointer");
    if (col_ptr_bytes != 4 && col_ptr_bytes != 8) fail(BFSX_ERR_ARG, "col_ptr_bytes must be 4 or 8");
    if (n_cols == 0 || n_cols > NA_MAX)
      fail(BFSX_ERR_ARG, "n_cols must be in [1, BFSX_NOALLOC_MAX_VERTICES = " +
                             std::to_string(NA_MAX) + "]");
    if (n_rows > n_cols) fail(BFSX_ERR_ARG, "n_rows > n_cols");
    if (root >= n_cols) fail(BFSX_ERR_ARG, "root out of range");
    if (col_ptr_bytes == 4 && nnz > 0xFFFFFFFFull) fail(BFSX_ERR_ARG, "nnz does not fit 4-byte col_ptr");
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) fail(BFSX_ERR_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n_dev || device >= NA_MAX_DEVICES) fail(BFSX_ERR_ARG, "device index out of range");
    BFSX_HIP(hipGetDevice(&old));
    BFSX_HIP(hipSetDevice(device));
    int cus = 0;
    BFSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    uint8_t *base = static_cast<uint8_t *>(na_scratch(device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t nl =
        col_ptr_bytes == 8
            ? na_search<uint64_t>(base, s, static_cast<const uint64_t *>(col_ptr), row_idx, nnz,
                                  n_cols, n_rows, root, levels, cus)
            : na_search<uint32_t>(base, s, static_cast<const uint32_t *>(col_ptr), row_idx, nnz,
                                  n_cols, n_rows, root, levels, cus);
    if (n_levels) *n_levels = nl;
    (void)hipSetDevice(old);
    na_error.clear();
    return BFSX_OK;
  } catch (const Error &e) {
    if (old >= 0) (void)hipSetDevice(old);
    na_error = e.msg;
    return e.status;
  } catch (...) {
    if (old >= 0) (void)hipSetDevice(old);
    na_error = "unknown failure";
    return BFSX_ERR_DEVICE;
  }
}
