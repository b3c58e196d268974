// Learned join-cardinality estimator: a 40 -> 64 -> 64 -> 1 ReLU MLP with BF16 operands and FP32
// accumulation on MFMA (v_mfma_f32_16x16x16_bf16), trained on the GPU with hand-written forward
// and backward kernels and Adam.  It predicts log2(1 + output rows) of a join from features the
// host computes; the prediction only presizes buffers (qf_create_binding_table_presized falls
// back exactly on overflow), so results never depend on it.
//
// Parameters (FP32 master copy, one buffer): W1[64][48] (inputs 40..47 are zero padding),
// b1[64], W2[64][64], b2[64], w3[64], b3.  GEMM operands are rounded to BF16 (round to nearest
// even) where they enter an MFMA: inputs x, W1, hidden h1, W2 in the forward pass; dz2, h1, W2,
// dz1 and x in the backward pass.  Everything else is FP32.
#include "qf_internal.h"

namespace qf {

constexpr int MLP_IN = 40;
constexpr int MLP_K1 = 48;  // inputs padded to three 16-wide MFMA steps
constexpr int MLP_H = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_K1;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H;
constexpr int MLP_PARAMS = OFF_B3 + 1;
constexpr int MLP_TRAIN_ROWS = 128;  // rows per training workgroup

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One 16x16 output tile on a wave: acc += A(16 x K) * B(K x 16), K a multiple of 16, operands
// BF16 in LDS with A(m, k) = a[m * sam + k * sak] and B(k, n) = b[k * sbk + n * sbn].  Lane l
// supplies A row l % 16 and B column l % 16 at k = 4 * (l / 16) + 0..3 of each 16-wide step and
// receives D rows 4 * (l / 16) + 0..3 of column l % 16.
__device__ inline f32x4 mfma_tile(const uint16_t* a, int sam, int sak, const uint16_t* b, int sbk, int sbn, int K,
                                  f32x4 acc, int lane) {
  const int r = lane & 15, kg = (lane >> 4) * 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      av[j] = (short)a[r * sam + (k0 + kg + j) * sak];
      bv[j] = (short)b[(k0 + kg + j) * sbk + r * sbn];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// Weights to BF16 in LDS: w1[64][48], w2[64][64].
__device__ inline void stage_weights(const float* p, uint16_t* w1, uint16_t* w2) {
  for (int i = threadIdx.x; i < MLP_H * MLP_K1; i += blockDim.x) w1[i] = f2bf(p[OFF_W1 + i]);
  for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) w2[i] = f2bf(p[OFF_W2 + i]);
}

// Forward of `rows` rows (a multiple of 16) starting at LDS row 0: xs[rows][48] BF16 in, h1
// BF16 and the ReLU masks of both layers, h2 FP32 and the predictions out.  Waves take 16-row
// tiles.
__device__ void forward_rows(const float* p, const uint16_t* w1, const uint16_t* w2, const uint16_t* xs, int rows,
                             uint16_t* h1, uint64_t* mask1, float* h2, float* pred, int lane, int wave, int nwaves) {
  const int col = lane & 15, rg = (lane >> 4) * 4;
  for (int t = wave; t < rows / 16; t += nwaves) {
    const int r0 = t * 16;
    for (int nt = 0; nt < MLP_H / 16; ++nt) {
      f32x4 acc = {0, 0, 0, 0};
      acc = mfma_tile(xs + r0 * MLP_K1, MLP_K1, 1, w1 + nt * 16 * MLP_K1, 1, MLP_K1, MLP_K1, acc, lane);
      const int n = nt * 16 + col;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float z = acc[j] + p[OFF_B1 + n];
        const int row = r0 + rg + j;
        h1[row * MLP_H + n] = f2bf(z > 0.f ? z : 0.f);
        if (mask1 && z > 0.f) atomicOr((unsigned long long*)&mask1[row], 1ull << n);
      }
    }
  }
  __syncthreads();
  for (int t = wave; t < rows / 16; t += nwaves) {
    const int r0 = t * 16;
    for (int nt = 0; nt < MLP_H / 16; ++nt) {
      f32x4 acc = {0, 0, 0, 0};
      acc = mfma_tile(h1 + r0 * MLP_H, MLP_H, 1, w2 + nt * 16 * MLP_H, 1, MLP_H, MLP_H, acc, lane);
      const int n = nt * 16 + col;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float z = acc[j] + p[OFF_B2 + n];
        h2[(r0 + rg + j) * MLP_H + n] = z > 0.f ? z : 0.f;
      }
    }
  }
  __syncthreads();
  for (int row = threadIdx.x; row < rows; row += blockDim.x) {
    float s = p[OFF_B3];
    for (int k = 0; k < MLP_H; ++k) s += h2[row * MLP_H + k] * p[OFF_W3 + k];
    pred[row] = s;
  }
  __syncthreads();
}

__device__ inline void stage_inputs(const float* X, int64_t row0, int64_t n, int rows, uint16_t* xs) {
  for (int i = threadIdx.x; i < rows * MLP_K1; i += blockDim.x) {
    const int r = i / MLP_K1, k = i % MLP_K1;
    xs[i] = (row0 + r < n && k < MLP_IN) ? f2bf(X[(row0 + r) * MLP_IN + k]) : (uint16_t)0;
  }
}

// Batched inference: 64 rows per 256-thread workgroup.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* p, const float* X, int64_t n, float* out) {
  __shared__ uint16_t w1[MLP_H * MLP_K1], w2[MLP_H * MLP_H], xs[64 * MLP_K1], h1[64 * MLP_H];
  __shared__ float h2[64 * MLP_H], pred[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  stage_weights(p, w1, w2);
  for (int64_t row0 = (int64_t)blockIdx.x * 64; row0 < n; row0 += (int64_t)gridDim.x * 64) {
    stage_inputs(X, row0, n, 64, xs);
    __syncthreads();
    forward_rows(p, w1, w2, xs, 64, h1, nullptr, h2, pred, lane, wave, 4);
    for (int r = threadIdx.x; r < 64; r += blockDim.x)
      if (row0 + r < n) out[row0 + r] = pred[r];
    __syncthreads();
  }
}

// Gradients of the mean squared error over `total` rows, for this workgroup's 128 rows, added
// into grads (FP32 atomics); loss_sum gets the squared errors.
__global__ __launch_bounds__(256) void k_mlp_grad(const float* p, const float* X, const float* y, int64_t n,
                                                  int64_t total, float* grads, float* loss_sum) {
  extern __shared__ uint8_t smem[];
  constexpr int R = MLP_TRAIN_ROWS;
  uint16_t* w1 = reinterpret_cast<uint16_t*>(smem);
  uint16_t* w2 = w1 + MLP_H * MLP_K1;
  uint16_t* xs = w2 + MLP_H * MLP_H;
  uint16_t* h1 = xs + R * MLP_K1;
  uint16_t* dz2 = h1 + R * MLP_H;
  uint16_t* dz1 = dz2 + R * MLP_H;
  float* h2 = reinterpret_cast<float*>(dz1 + R * MLP_H);
  float* pred = h2 + R * MLP_H;
  float* dpred = pred + R;
  uint64_t* mask1 = reinterpret_cast<uint64_t*>(dpred + R);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  stage_weights(p, w1, w2);
  stage_inputs(X, row0, n, R, xs);
  for (int r = threadIdx.x; r < R; r += blockDim.x) mask1[r] = 0;
  __syncthreads();
  forward_rows(p, w1, w2, xs, R, h1, mask1, h2, pred, lane, wave, 4);
  // Output layer.
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    const bool valid = row0 + r < n;
    const float e = valid ? pred[r] - y[row0 + r] : 0.f;
    dpred[r] = 2.f * e / (float)total;
    if (valid) atomicAdd(loss_sum, e * e);
  }
  __syncthreads();
  for (int k = threadIdx.x; k <= MLP_H; k += blockDim.x) {
    float g = 0.f;
    for (int r = 0; r < R; ++r) g += dpred[r] * (k < MLP_H ? h2[r * MLP_H + k] : 1.f);
    atomicAdd(&grads[k < MLP_H ? OFF_W3 + k : OFF_B3], g);
  }
  // dz2 = dpred * w3 * relu'(z2).
  for (int i = threadIdx.x; i < R * MLP_H; i += blockDim.x) {
    const int r = i / MLP_H, k = i % MLP_H;
    dz2[i] = f2bf(h2[i] > 0.f ? dpred[r] * p[OFF_W3 + k] : 0.f);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) {
    float g = 0.f;
    for (int r = 0; r < R; ++r) g += bf2f(dz2[r * MLP_H + k]);
    atomicAdd(&grads[OFF_B2 + k], g);
  }
  const int col = lane & 15, rg = (lane >> 4) * 4;
  // dW2[n][k] = sum_r dz2[r][n] h1[r][k]: 4 x 4 tiles over the waves.
  for (int tile = wave; tile < 16; tile += 4) {
    const int mt = tile / 4, nt = tile % 4;
    f32x4 acc = {0, 0, 0, 0};
    acc = mfma_tile(dz2 + mt * 16, 1, MLP_H, h1 + nt * 16, MLP_H, 1, R, acc, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) atomicAdd(&grads[OFF_W2 + (mt * 16 + rg + j) * MLP_H + nt * 16 + col], acc[j]);
  }
  // dh1 = dz2 W2; dz1 = dh1 * relu'(z1).
  for (int t = wave; t < R / 16; t += 4) {
    for (int nt = 0; nt < MLP_H / 16; ++nt) {
      f32x4 acc = {0, 0, 0, 0};
      acc = mfma_tile(dz2 + t * 16 * MLP_H, MLP_H, 1, w2 + nt * 16, MLP_H, 1, MLP_H, acc, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = t * 16 + rg + j, k = nt * 16 + col;
        dz1[r * MLP_H + k] = f2bf(((mask1[r] >> k) & 1ull) ? acc[j] : 0.f);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) {
    float g = 0.f;
    for (int r = 0; r < R; ++r) g += bf2f(dz1[r * MLP_H + k]);
    atomicAdd(&grads[OFF_B1 + k], g);
  }
  // dW1[n][k] = sum_r dz1[r][n] x[r][k]: 4 x 3 tiles.
  for (int tile = wave; tile < 12; tile += 4) {
    const int mt = tile / 3, nt = tile % 3;
    f32x4 acc = {0, 0, 0, 0};
    acc = mfma_tile(dz1 + mt * 16, 1, MLP_H, xs + nt * 16, MLP_K1, 1, R, acc, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = nt * 16 + col;
      if (k < MLP_IN) atomicAdd(&grads[OFF_W1 + (mt * 16 + rg + j) * MLP_K1 + k], acc[j]);
    }
  }
}

// Adam (Kingma & Ba) on the FP32 master parameters.
__global__ void k_mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.f - b1) * gi;
  v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
  p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

size_t grad_lds_bytes() {
  constexpr int R = MLP_TRAIN_ROWS;
  return (size_t)(MLP_H * MLP_K1 + MLP_H * MLP_H + R * MLP_K1 + 3 * R * MLP_H) * 2 +
         (size_t)(R * MLP_H + 2 * R) * 4 + (size_t)R * 8;
}

}  // namespace qf

using namespace qf;

struct qf_mlp {
  DevBuf params, grads, m, v, loss;
  int64_t step = 0;
};

extern "C" {

int qf_mlp_num_params(void) { return MLP_PARAMS; }

int qf_mlp_create(qf_context* c, const float* h_params, qf_mlp** out) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(out && h_params, "null argument");
    std::unique_ptr<qf_mlp> m(new qf_mlp);
    m->params = DevBuf(ctx, MLP_PARAMS * 4);
    m->grads = DevBuf(ctx, MLP_PARAMS * 4);
    m->m = DevBuf(ctx, MLP_PARAMS * 4);
    m->v = DevBuf(ctx, MLP_PARAMS * 4);
    m->loss = DevBuf(ctx, 4);
    copy_h2d(ctx, m->params.get(), h_params, MLP_PARAMS * 4);
    fill_i32(ctx, m->m.as<int32_t>(), 0, MLP_PARAMS);
    fill_i32(ctx, m->v.as<int32_t>(), 0, MLP_PARAMS);
    sync(ctx);
    *out = m.release();
  });
}

void qf_mlp_destroy(qf_mlp* m) { delete m; }

int qf_mlp_get_params(qf_context* c, qf_mlp* m, float* h_params) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(m && h_params, "null argument");
    copy_d2h(ctx, h_params, m->params.get(), MLP_PARAMS * 4);
  });
}

int qf_mlp_forward(qf_context* c, qf_mlp* m, const float* d_x, int64_t n, float* d_out) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(m && (n == 0 || (d_x && d_out)) && n >= 0, "bad argument");
    if (n == 0) return;
    const int blocks = (int)std::min<int64_t>(4 * ctx->num_cus, ceil_div(n, 64));
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, ctx->stream, m->params.as<float>(), d_x, n, d_out);
    QF_CHECK_HIP(hipGetLastError());
  });
}

int qf_mlp_gradients(qf_context* c, qf_mlp* m, const float* d_x, const float* d_y, int64_t n, float* h_grads,
                     float* h_loss) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(m && d_x && d_y && n > 0, "bad argument");
    fill_i32(ctx, m->grads.as<int32_t>(), 0, MLP_PARAMS);
    fill_i32(ctx, m->loss.as<int32_t>(), 0, 1);
    const size_t lds = grad_lds_bytes();
    QF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_grad),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)ceil_div(n, MLP_TRAIN_ROWS)), dim3(256), lds, ctx->stream,
                       m->params.as<float>(), d_x, d_y, n, n, m->grads.as<float>(), m->loss.as<float>());
    QF_CHECK_HIP(hipGetLastError());
    if (h_grads) copy_d2h(ctx, h_grads, m->grads.get(), MLP_PARAMS * 4);
    if (h_loss) {
      copy_d2h(ctx, h_loss, m->loss.get(), 4);
      *h_loss /= (float)n;
    }
  });
}

int qf_mlp_train_step(qf_context* c, qf_mlp* m, const float* d_x, const float* d_y, int64_t n, float lr,
                      float* h_loss) {
  const int st = qf_mlp_gradients(c, m, d_x, d_y, n, nullptr, h_loss);
  if (st != QF_OK) return st;
  return guarded(c, [&](Context* ctx) {
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)ceil_div(MLP_PARAMS, 256)), dim3(256), 0, ctx->stream,
                       m->params.as<float>(), m->grads.as<float>(), m->m.as<float>(), m->v.as<float>(), MLP_PARAMS, lr,
                       b1, b2, 1e-8f, c1, c2);
    QF_CHECK_HIP(hipGetLastError());
  });
}

}  // extern "C"
