// Learned per-level scheduler: an FP32 MLP 8 -> 32 -> 32 -> 4 (ReLU) that,
// before each level, predicts log2 of the device time of each step kind
// (LLB, GLB, SLB, PULL) relative to LLB from frontier statistics; the level
// runs the step with the smallest prediction.  It only chooses the
// schedule: every step kind computes the same exact levels.  The inputs and
// the inference are in bfsx_mlp.hpp.
// Parameter layout (1476 floats): W1[32][8], b1[32], W2[32][32], b2[32],
// W3[4][32], b3[4], row-major.
//
// Training is full-batch Adam on the GPU: a per-sample forward/backward
// kernel stores activations and deltas column-major, a gradient kernel
// reduces each parameter row over all samples in a fixed order (one block
// per row, deterministic), and an Adam kernel updates the parameters.
#include "bfsx_mlp.hpp"

namespace bfsx {

// ---------------------------------------------------------------------------
// inference before the first level (later levels decide in k_level_end)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_decide(LevelCtl *ctl, const float *Wg, uint32_t idx,
                                               uint64_t n_cols, uint64_t nnz, uint64_t n_iso,
                                               uint32_t allow_mask, int first) {
  decide_wave(ctl, Wg, idx, n_cols, nnz, n_iso, allow_mask, first);
}

void decide(bfsx_ctx *ctx, uint32_t idx, bool first, uint32_t allow_mask) {
  const uint64_t n_iso = (ctx->gflags & BFSX_GRAPH_SYMMETRIC) && ctx->n_rows == ctx->n_cols
                             ? ctx->n_isolated
                             : 0;
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, ctx->stream, ctx->ctl.as<LevelCtl>(),
                     ctx->sched_w.as<float>(), idx, (uint64_t)ctx->n_cols, (uint64_t)ctx->nnz,
                     n_iso, allow_mask, first ? 1 : 0);
  BFSX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// batch forward (thread per sample)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void forward_one(const float *W, const float *x, float *a1, float *a2,
                                            float *y) {
  for (int j = 0; j < H1; j++) {
    float s = W[OFF_B1 + j];
    for (int k = 0; k < IN; k++) s = fmaf(W[OFF_W1 + j * IN + k], x[k], s);
    a1[j] = s > 0.f ? s : 0.f;
  }
  for (int j = 0; j < H2; j++) {
    float s = W[OFF_B2 + j];
    for (int k = 0; k < H1; k++) s = fmaf(W[OFF_W2 + j * H1 + k], a1[k], s);
    a2[j] = s > 0.f ? s : 0.f;
  }
  for (int j = 0; j < OUT; j++) {
    float s = W[OFF_B3 + j];
    for (int k = 0; k < H2; k++) s = fmaf(W[OFF_W3 + j * H2 + k], a2[k], s);
    y[j] = s;
  }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *X, uint32_t n, const float *Wg,
                                                     float *Y) {
  __shared__ float W[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) W[i] = Wg[i];
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[IN], a1[H1], a2[H2], y[OUT];
  for (int k = 0; k < IN; k++) x[k] = X[(size_t)i * IN + k];
  forward_one(W, x, a1, a2, y);
  for (int j = 0; j < OUT; j++) Y[(size_t)i * OUT + j] = y[j];
}

// ---------------------------------------------------------------------------
// training: forward + backward per sample, deterministic gradient reduction
// ---------------------------------------------------------------------------
// Loss = sum_ij M_ij (y_ij - t_ij)^2 * inv_count, inv_count = 1 / sum M.
// Activations/deltas are stored column-major: A1[j * n + i].
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float *X, const float *T,
                                                     const float *M, uint32_t n,
                                                     float inv_count, const float *Wg,
                                                     float *A1, float *A2, float *D1, float *D2,
                                                     float *D3, float *L) {
  __shared__ float W[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) W[i] = Wg[i];
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[IN], a1[H1], a2[H2], y[OUT], d3[OUT], d2[H2];
  for (int k = 0; k < IN; k++) x[k] = X[(size_t)i * IN + k];
  forward_one(W, x, a1, a2, y);
  float loss = 0.f;
  for (int j = 0; j < OUT; j++) {
    const float m = M[(size_t)i * OUT + j];
    const float e = y[j] - T[(size_t)i * OUT + j];
    loss += m * e * e * inv_count;
    d3[j] = 2.f * m * e * inv_count;
    D3[(size_t)j * n + i] = d3[j];
  }
  L[i] = loss;
  for (int k = 0; k < H2; k++) {
    float s = 0.f;
    for (int j = 0; j < OUT; j++) s = fmaf(W[OFF_W3 + j * H2 + k], d3[j], s);
    d2[k] = a2[k] > 0.f ? s : 0.f;
    D2[(size_t)k * n + i] = d2[k];
    A2[(size_t)k * n + i] = a2[k];
  }
  for (int k = 0; k < H1; k++) {
    float s = 0.f;
    for (int j = 0; j < H2; j++) s = fmaf(W[OFF_W2 + j * H1 + k], d2[j], s);
    D1[(size_t)k * n + i] = a1[k] > 0.f ? s : 0.f;
    A1[(size_t)k * n + i] = a1[k];
  }
}

// Fixed-order block reduction of one value per thread.
__device__ __forceinline__ float block_reduce_256(float v, float *s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  const float r = s[0];
  __syncthreads();
  return r;
}

// Block b < 32: layer-1 row b; 32 <= b < 64: layer-2 row b-32; 64 <= b < 68:
// layer-3 row b-64; block 68: the loss.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *X, const float *A1,
                                                  const float *A2, const float *D1,
                                                  const float *D2, const float *D3,
                                                  const float *L, uint32_t n, float *G,
                                                  float *loss_out) {
  __shared__ float s[256];
  const int b = blockIdx.x;
  float acc[H1 + 1];
  for (int k = 0; k <= H1; k++) acc[k] = 0.f;
  if (b < H1) {
    const int j = b;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
      const float d = D1[(size_t)j * n + i];
      for (int k = 0; k < IN; k++) acc[k] = fmaf(d, X[(size_t)i * IN + k], acc[k]);
      acc[IN] += d;
    }
    for (int k = 0; k <= IN; k++) {
      const float r = block_reduce_256(acc[k], s);
      if (threadIdx.x == 0) G[k < IN ? OFF_W1 + j * IN + k : OFF_B1 + j] = r;
    }
  } else if (b < H1 + H2) {
    const int j = b - H1;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
      const float d = D2[(size_t)j * n + i];
      for (int k = 0; k < H1; k++) acc[k] = fmaf(d, A1[(size_t)k * n + i], acc[k]);
      acc[H1] += d;
    }
    for (int k = 0; k <= H1; k++) {
      const float r = block_reduce_256(acc[k], s);
      if (threadIdx.x == 0) G[k < H1 ? OFF_W2 + j * H1 + k : OFF_B2 + j] = r;
    }
  } else if (b < H1 + H2 + OUT) {
    const int j = b - H1 - H2;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
      const float d = D3[(size_t)j * n + i];
      for (int k = 0; k < H2; k++) acc[k] = fmaf(d, A2[(size_t)k * n + i], acc[k]);
      acc[H2] += d;
    }
    for (int k = 0; k <= H2; k++) {
      const float r = block_reduce_256(acc[k], s);
      if (threadIdx.x == 0) G[k < H2 ? OFF_W3 + j * H2 + k : OFF_B3 + j] = r;
    }
  } else {
    float v = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += 256) v += L[i];
    const float r = block_reduce_256(v, s);
    if (threadIdx.x == 0) *loss_out = r;
  }
}

__global__ void k_adam(float *W, const float *G, float *m, float *v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  W[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct TrainBufs {
  DevBuf X, T, M, W, G, A1, A2, D1, D2, D3, L, m, v, loss;
};

static void upload_batch(bfsx_ctx *ctx, TrainBufs &b, const float *X, const float *T,
                         const float *M, uint32_t n, const float *W) {
  hipStream_t s = ctx->stream;
  ensure(b.X, sizeof(float) * IN * (size_t)n, s);
  ensure(b.W, sizeof(float) * NPARAM, s);
  BFSX_HIP(hipMemcpyAsync(b.X.p, X, sizeof(float) * IN * (size_t)n, hipMemcpyHostToDevice, s));
  BFSX_HIP(hipMemcpyAsync(b.W.p, W, sizeof(float) * NPARAM, hipMemcpyHostToDevice, s));
  if (T) {
    ensure(b.T, sizeof(float) * OUT * (size_t)n, s);
    ensure(b.M, sizeof(float) * OUT * (size_t)n, s);
    BFSX_HIP(hipMemcpyAsync(b.T.p, T, sizeof(float) * OUT * (size_t)n, hipMemcpyHostToDevice, s));
    BFSX_HIP(hipMemcpyAsync(b.M.p, M, sizeof(float) * OUT * (size_t)n, hipMemcpyHostToDevice, s));
    for (DevBuf *d : {&b.A1, &b.A2, &b.D1, &b.D2})
      ensure(*d, sizeof(float) * 32 * (size_t)n, s);
    ensure(b.D3, sizeof(float) * OUT * (size_t)n, s);
    ensure(b.L, sizeof(float) * (size_t)n, s);
    ensure(b.G, sizeof(float) * NPARAM, s);
  }
}

static void release_batch(TrainBufs &b) {
  for (DevBuf *d : {&b.X, &b.T, &b.M, &b.W, &b.G, &b.A1, &b.A2, &b.D1, &b.D2, &b.D3, &b.L, &b.m,
                    &b.v, &b.loss})
    release(*d);
}

static float mask_inv_count(const float *M, uint32_t n) {
  double c = 0;
  for (size_t i = 0; i < (size_t)n * OUT; i++) c += M[i];
  if (c <= 0) fail(BFSX_ERR_ARG, "mask selects no target");
  return (float)(1.0 / c);
}

static void grad_pass(bfsx_ctx *ctx, TrainBufs &b, uint32_t n, float inv, float *d_loss) {
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((n + 255) / 256), dim3(256), 0, s, b.X.as<float>(),
                     b.T.as<float>(), b.M.as<float>(), n, inv, b.W.as<float>(), b.A1.as<float>(),
                     b.A2.as<float>(), b.D1.as<float>(), b.D2.as<float>(), b.D3.as<float>(),
                     b.L.as<float>());
  hipLaunchKernelGGL(k_mlp_grad, dim3(H1 + H2 + OUT + 1), dim3(256), 0, s, b.X.as<float>(),
                     b.A1.as<float>(), b.A2.as<float>(), b.D1.as<float>(), b.D2.as<float>(),
                     b.D3.as<float>(), b.L.as<float>(), n, b.G.as<float>(), d_loss);
  BFSX_HIP(hipGetLastError());
}

void mlp_forward(bfsx_ctx *ctx, const float *X, uint32_t n, const float *W, float *Y) {
  TrainBufs b;
  try {
    upload_batch(ctx, b, X, nullptr, nullptr, n, W);
    DevBuf y;
    ensure(y, sizeof(float) * OUT * (size_t)n, ctx->stream);
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       b.X.as<float>(), n, b.W.as<float>(), y.as<float>());
    BFSX_HIP(hipGetLastError());
    BFSX_HIP(hipMemcpyAsync(Y, y.p, sizeof(float) * OUT * (size_t)n, hipMemcpyDeviceToHost,
                            ctx->stream));
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
    release(y);
  } catch (...) {
    release_batch(b);
    throw;
  }
  release_batch(b);
}

void mlp_loss_grad(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
                   const float *W, float *grad, float *loss) {
  TrainBufs b;
  try {
    const float inv = mask_inv_count(M, n);
    upload_batch(ctx, b, X, T, M, n, W);
    ensure(b.loss, sizeof(float), ctx->stream);
    grad_pass(ctx, b, n, inv, b.loss.as<float>());
    BFSX_HIP(hipMemcpyAsync(grad, b.G.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost,
                            ctx->stream));
    BFSX_HIP(hipMemcpyAsync(loss, b.loss.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    BFSX_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    release_batch(b);
    throw;
  }
  release_batch(b);
}

void mlp_train(bfsx_ctx *ctx, const float *X, const float *T, const float *M, uint32_t n,
               float *W, uint32_t epochs, float lr, float *loss_curve) {
  TrainBufs b;
  try {
    const float inv = mask_inv_count(M, n);
    upload_batch(ctx, b, X, T, M, n, W);
    hipStream_t s = ctx->stream;
    ensure(b.m, sizeof(float) * NPARAM, s);
    ensure(b.v, sizeof(float) * NPARAM, s);
    ensure(b.loss, sizeof(float) * (epochs ? epochs : 1), s);
    BFSX_HIP(hipMemsetAsync(b.m.p, 0, sizeof(float) * NPARAM, s));
    BFSX_HIP(hipMemsetAsync(b.v.p, 0, sizeof(float) * NPARAM, s));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double p1 = 1.0, p2 = 1.0;
    for (uint32_t e = 0; e < epochs; e++) {
      grad_pass(ctx, b, n, inv, b.loss.as<float>() + e);
      p1 *= b1;
      p2 *= b2;
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, b.W.as<float>(),
                         b.G.as<float>(), b.m.as<float>(), b.v.as<float>(), lr, b1, b2, eps,
                         (float)(1.0 - p1), (float)(1.0 - p2));
      BFSX_HIP(hipGetLastError());
    }
    BFSX_HIP(hipMemcpyAsync(W, b.W.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
    if (loss_curve && epochs)
      BFSX_HIP(hipMemcpyAsync(loss_curve, b.loss.p, sizeof(float) * epochs,
                              hipMemcpyDeviceToHost, s));
    BFSX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    release_batch(b);
    throw;
  }
  release_batch(b);
}

}  // namespace bfsx
