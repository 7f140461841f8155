// Device side of the learned per-level scheduler: the network's layout and
// its inference, shared by the stand-alone decision kernel and the
// level-end kernel (which decides the next level in the same launch).
#pragma once
#include "bfsx_internal.hpp"

namespace bfsx {

constexpr int IN = 8, H1 = 32, H2 = 32, OUT = 4;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H1 * IN, OFF_W2 = OFF_B1 + H1,
              OFF_B2 = OFF_W2 + H2 * H1, OFF_W3 = OFF_B2 + H2, OFF_B3 = OFF_W3 + OUT * H2,
              NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == BFSX_SCHED_PARAMS, "parameter count");
static_assert(NPARAM % 4 == 0, "weights load as float4");

// log2((1 + a) / (1 + b)) for a, b >= 0
__device__ __forceinline__ float flog2ratio(double a, double b) {
  return (float)(log2(1.0 + (a > 0 ? a : 0.0)) - log2(1.0 + (b > 0 ? b : 0.0)));
}

// Inference by one work-group of 64 threads, in two parts: every thread
// stages the weights into LDS (float4 loads), then, after a barrier,
// decide_compute: the inputs of
// frontier buffer `idx` from the control block, the three layers, and
// ctl->choice = argmin over the allowed kinds (fixed LLB when an output is
// NaN); inputs and outputs are recorded in the control block.  Returns the
// choice in thread 0 (0 in the others).
//   x0 = log2((1+F)/(1+n))/32          x1 = log2((1+EF)/(1+nnz))/32
//   x2 = log2((1+UV)/(1+n))/32         x3 = log2((1+UE)/(1+nnz))/32
//   x4 = log2(1+EF/F)/16               x5 = log2(1+nnz/n)/16
//   x6 = log2(1+n)/32                  x7 = 1 if the last step was bottom-up
// F/EF: frontier vertices/edges, UV/UE: unvisited vertices/edges (isolated
// vertices count as visited).
__device__ __forceinline__ void decide_stage(const float *Wg, float4 *W4) {
  const float4 *G4 = reinterpret_cast<const float4 *>(Wg);
  for (int i = threadIdx.x; i < NPARAM / 4; i += 64) W4[i] = G4[i];
}

__device__ __forceinline__ uint32_t decide_compute(LevelCtl *ctl, const float4 *W4, uint32_t idx,
                                               uint64_t n_cols, uint64_t nnz, uint64_t n_iso,
                                               uint32_t allow_mask, int first) {
  __shared__ float x[IN], h1[H1], h2[H2], y[OUT];
  const float *W = reinterpret_cast<const float *>(W4);
  const int l = threadIdx.x;
  if (l < IN) {
    // lane l computes input l; every lane runs the same instructions
    // (x = log2((1+a)/(1+b))/d with a, b, d picked per lane)
    const double F = ctl->q_count[idx];
    const double EF = (double)ctl->q_edges[idx];
    const double vv = (double)ctl->visited_vertices + (double)n_iso;
    const double UV = vv < (double)n_cols ? (double)n_cols - vv : 0.0;
    const double ve = (double)ctl->visited_edges;
    const double UE = ve < (double)nnz ? (double)nnz - ve : 0.0;
    const double n = (double)n_cols, m = (double)nnz;
    const double a = l == 0 ? F : l == 1 ? EF : l == 2 ? UV : l == 3 ? UE
                   : l == 4 ? (F > 0 ? EF / F : 0.0) : l == 5 ? (n > 0 ? m / n : 0.0) : n;
    const double b = (l == 0 || l == 2) ? n : (l == 1 || l == 3) ? m : 0.0;
    float v = flog2ratio(a, b) / ((l == 4 || l == 5) ? 16.f : 32.f);
    if (l == 7) v = (!first && ctl->choice == BFSX_PULL) ? 1.f : 0.f;
    x[l] = v;
  }
  __syncthreads();
  if (l < H1) {
    float s = W[OFF_B1 + l];
    for (int k = 0; k < IN; k++) s = fmaf(W[OFF_W1 + l * IN + k], x[k], s);
    h1[l] = s > 0.f ? s : 0.f;
  }
  __syncthreads();
  if (l < H2) {
    float s = W[OFF_B2 + l];
    for (int k = 0; k < H1; k++) s = fmaf(W[OFF_W2 + l * H1 + k], h1[k], s);
    h2[l] = s > 0.f ? s : 0.f;
  }
  __syncthreads();
  if (l < OUT) {
    float s = W[OFF_B3 + l];
    for (int k = 0; k < H2; k++) s = fmaf(W[OFF_W3 + l * H2 + k], h2[k], s);
    y[l] = s;
  }
  __syncthreads();
  uint32_t picked = 0;
  if (l == 0) {
    uint32_t best = 0;
    float bv = INFINITY;
    bool ok = true;
    for (int j = 0; j < OUT; j++) {
      ctl->sched_out[j] = y[j];
      if (!(y[j] == y[j])) ok = false;  // NaN
      if (((allow_mask >> j) & 1u) && y[j] < bv) {
        bv = y[j];
        best = (uint32_t)j;
      }
    }
    for (int k = 0; k < IN; k++) ctl->sched_in[k] = x[k];
    picked = ok && bv < INFINITY ? best : (uint32_t)BFSX_LLB;
    ctl->choice = picked;
    ctl->decisions += 1;
  }
  return picked;
}

__device__ __forceinline__ void decide_wave(LevelCtl *ctl, const float *Wg, uint32_t idx,
                                            uint64_t n_cols, uint64_t nnz, uint64_t n_iso,
                                            uint32_t allow_mask, int first) {
  __shared__ float4 W4[NPARAM / 4];
  decide_stage(Wg, W4);
  __syncthreads();
  decide_compute(ctl, W4, idx, n_cols, nnz, n_iso, allow_mask, first);
}

}  // namespace bfsx
