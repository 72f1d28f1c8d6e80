// mlp.hip -- per-particle pressure predictor 8 -> 64 -> 64 -> 1 (SiLU) on the
// gfx950 matrix cores: BF16 operands, FP32 accumulation
// (v_mfma_f32_32x32x16_bf16), forward and backward, Adam in FP32.
//
// One wave evaluates a tile of 32 rows (building blocks in mlp_dev.h).
#include <hip/hip_runtime.h>
#include <string.h>

#include "mlp_dev.h"

namespace wsph {

// Per-wave LDS scratch: 32 rows x 64 columns of float (activations) plus
// a second buffer for pre-activations in training.
struct WaveLds {
  float a[32][65];
  float b[32][65];
};

// ---- parameter conversion ---------------------------------------------------
// params (fp32): W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3
__global__ void k_to_bf16(const float *__restrict__ p, __bf16 *__restrict__ w1,
                          __bf16 *__restrict__ w2, __bf16 *__restrict__ w3) {
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < WSPH_MLP_H * WSPH_MLP_H;
       k += blockDim.x * gridDim.x) {
    if (k < WSPH_MLP_H * WSPH_MLP_IN) w1[k] = (__bf16)p[WSPH_MLP_OFF_W1 + k];
    w2[k] = (__bf16)p[WSPH_MLP_OFF_W2 + k];
    if (k < WSPH_MLP_H) w3[k] = (__bf16)p[WSPH_MLP_OFF_W3 + k];
  }
}


// ---- inference on a feature matrix -------------------------------------------
// 64 rows per wave and trip, register-resident (infer64).
constexpr int kMlpWaves = 4;

__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_forward(Weights W, const float *__restrict__ x,
                                                                int n, float *__restrict__ y) {
  __shared__ NetLds L;
  stage_net(L, W);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long waves = (long)gridDim.x * kMlpWaves;
  for (long r0 = ((long)blockIdx.x * kMlpWaves + (threadIdx.x >> 6)) * 64; r0 < n; r0 += waves * 64) {
    const long row = r0 + lane;
    float f[WSPH_MLP_IN];
#pragma unroll
    for (int k = 0; k < WSPH_MLP_IN; ++k) f[k] = row < n ? x[row * WSPH_MLP_IN + k] : 0.0f;
    const float v = infer64(L, f);
    if (row < n) y[row] = v;
  }
}

// ---- training: forward + backward of one batch -----------------------------------
// Per-block partial gradients (deterministic: reduced in block order).
constexpr int kBwdWaves = 2;  // LDS: activations, pre-activations and a gradient copy per wave
__global__ __launch_bounds__(64 * kBwdWaves) void k_mlp_backward(Weights W, const float *__restrict__ x,
                                                                 const float *__restrict__ t, int n,
                                                                 float inv_batch,
                                                                 float *__restrict__ part,
                                                                 float *__restrict__ sse_part) {
  __shared__ WaveLds lds[kBwdWaves];
  __shared__ float z1s[kBwdWaves][32 * 64];
  __shared__ float z2s[kBwdWaves][32 * 64];
  __shared__ float gacc[kBwdWaves][WSPH_MLP_NPARAM];
  __shared__ float dys[kBwdWaves][32];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  WaveLds &L = lds[wave];
  float *G = gacc[wave];
  for (int k = lane; k < WSPH_MLP_NPARAM; k += 64) G[k] = 0.0f;
  float sse = 0.0f;
  const int tiles = (n + 31) / 32;
  // fp32 accumulators of the weight gradients, in MFMA D layout
  f32x16 g2[2][2] = {};  // dW2: [M tile n][N tile k]
  f32x16 g1[2] = {};     // dW1: [M tile n], N = inputs (8 of 32 columns used)
  __syncthreads();
  for (int tt0 = blockIdx.x * kBwdWaves; tt0 < tiles; tt0 += gridDim.x * kBwdWaves) {
    const int tt = tt0 + wave;
    const bool live = tt < tiles;
    const int tile = live ? tt : tiles - 1;
    const int rows = live ? min(32, n - 32 * tile) : 0;
    const float *xt = x + (long)32 * tile * WSPH_MLP_IN;
    float y[16];
    forward_tile(W, xt, rows, L.a, z1s[wave], z2s[wave], y);
    // dL/dy = (y - t) / B on the lanes that hold column 0
    if ((lane & 31) == 0) {
      for (int i = 0; i < 16; ++i) {
        const int r = drow(lane, i);
        float dy = 0.0f;
        if (r < rows) {
          const float e = y[i] - t[(long)32 * tile + r];
          sse += e * e;
          dy = e * inv_batch;
        }
        dys[wave][r] = dy;
      }
    }
    __syncthreads();
    // layer 3: dW3[k] = sum_r dy_r h2[r][k]; db3 = sum_r dy_r; dz2 = dy w3 silu'(z2)
    for (int k = lane; k < 64; k += 64) {
      float s = 0.0f;
      for (int r = 0; r < 32; ++r) s += dys[wave][r] * L.a[r][k];
      G[WSPH_MLP_OFF_W3 + k] += s;
    }
    if (lane == 0) {
      float s = 0.0f;
      for (int r = 0; r < 32; ++r) s += dys[wave][r];
      G[WSPH_MLP_OFF_B3] += s;
    }
    __syncthreads();
    for (int e = lane; e < 32 * 64; e += 64) {
      const int r = e >> 6, k = e & 63;
      const float w3 = (float)W.w3[k];
      L.b[r][k] = dys[wave][r] * w3 * dsilu(z2s[wave][e]);  // dz2
    }
    __syncthreads();
    // db2[k] = sum_r dz2[r][k]
    for (int k = lane; k < 64; k += 64) {
      float s = 0.0f;
      for (int r = 0; r < 32; ++r) s += L.b[r][k];
      G[WSPH_MLP_OFF_B2 + k] += s;
    }
    // dW2[n][k] += sum_r dz2[r][n] h1[r][k]: A = dz2^T (M = n, K = r), B = h1 (K = r, N = k).
    // h1 = silu(z1) is rebuilt from z1s into L.a (L.a holds h2 now).
    __syncthreads();
    for (int e = lane; e < 32 * 64; e += 64) L.a[e >> 6][e & 63] = silu(z1s[wave][e]);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a;
        {
          const int nn = 32 * mt + (lane & 31), r0 = 16 * ks + 8 * (lane >> 5);
#pragma unroll
          for (int j = 0; j < 8; ++j) a[j] = (__bf16)L.b[r0 + j][nn];
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          bf16x8 bb;
          const int kk = 32 * nt + (lane & 31), r0 = 16 * ks + 8 * (lane >> 5);
#pragma unroll
          for (int j = 0; j < 8; ++j) bb[j] = (__bf16)L.a[r0 + j][kk];
          g2[mt][nt] = mfma(a, bb, g2[mt][nt]);
        }
      }
    // dh1[r][k] = sum_n dz2[r][n] W2[n][k]: A = dz2 (M = r, K = n), B = W2 (K = n, N = k)
    bf16x8 ad[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) ad[ks] = lds_a(L.b, lane, ks);
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      f32x16 c = {};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        bf16x8 bb;
        const int kk = 32 * nt + (lane & 31), n0 = 16 * ks + 8 * (lane >> 5);
#pragma unroll
        for (int j = 0; j < 8; ++j) bb[j] = W.w2[(n0 + j) * WSPH_MLP_H + kk];
        c = mfma(ad[ks], bb, c);
      }
      const int kk = 32 * nt + (lane & 31);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int r = drow(lane, i);
        L.b[r][kk] = r < rows ? c[i] * dsilu(z1s[wave][r * 64 + kk]) : 0.0f;  // dz1
      }
    }
    __syncthreads();
    // db1[n] = sum_r dz1[r][n]
    for (int k = lane; k < 64; k += 64) {
      float s = 0.0f;
      for (int r = 0; r < 32; ++r) s += L.b[r][k];
      G[WSPH_MLP_OFF_B1 + k] += s;
    }
    // dW1[n][k] += sum_r dz1[r][n] x[r][k]: A = dz1^T (M = n, K = r), B = x (K = r, N = k)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a, bb;
        const int nn = 32 * mt + (lane & 31), r0 = 16 * ks + 8 * (lane >> 5);
        const int kk = lane & 31;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int r = r0 + j;
          a[j] = (__bf16)L.b[r][nn];
          bb[j] = (kk < WSPH_MLP_IN && r < rows) ? (__bf16)xt[r * WSPH_MLP_IN + kk] : (__bf16)0.0f;
        }
        g1[mt] = mfma(a, bb, g1[mt]);
      }
    __syncthreads();
  }
  // accumulators -> per-wave gradient, then per-block partial
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int nn = 32 * mt + drow(lane, i);
      const int kk = lane & 31;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) G[WSPH_MLP_OFF_W2 + nn * WSPH_MLP_H + 32 * nt + kk] += g2[mt][nt][i];
      if (kk < WSPH_MLP_IN) G[WSPH_MLP_OFF_W1 + nn * WSPH_MLP_IN + kk] += g1[mt][i];
    }
  for (int o = 32; o > 0; o >>= 1) sse += __shfl_xor(sse, o, 64);
  __syncthreads();
  for (int k = threadIdx.x; k < WSPH_MLP_NPARAM; k += blockDim.x) {
    float s = 0.0f;
    for (int w = 0; w < kBwdWaves; ++w) s += gacc[w][k];
    part[(long)blockIdx.x * WSPH_MLP_NPARAM + k] = s;
  }
  if (lane == 0) dys[wave][0] = sse;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kBwdWaves; ++w) s += dys[w][0];
    sse_part[blockIdx.x] = s;
  }
}

// gradient = sum of block partials (block order), then one Adam step
__global__ void k_adam(float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2,
                       const float *__restrict__ part, int nparts, float *__restrict__ grad_out,
                       const float *__restrict__ sse_part, float *__restrict__ loss_out,
                       float inv_batch, float lr, float beta1, float beta2, float eps, float bc1,
                       float bc2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0 && loss_out) {
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += sse_part[b];
    *loss_out = 0.5f * s * inv_batch;
  }
  if (k >= WSPH_MLP_NPARAM) return;
  float g = 0.0f;
  for (int b = 0; b < nparts; ++b) g += part[(long)b * WSPH_MLP_NPARAM + k];
  if (grad_out) grad_out[k] = g;
  if (lr <= 0.0f) return;
  const float a = beta1 * m1[k] + (1.0f - beta1) * g;
  const float v = beta2 * m2[k] + (1.0f - beta2) * g * g;
  m1[k] = a;
  m2[k] = v;
  params[k] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

// ---- host side -----------------------------------------------------------------
hipError_t mlp_alloc(MlpDev &m, int max_batch) {
  memset(&m, 0, sizeof(m));
  m.max_batch = max_batch;
  hipError_t e;
  if ((e = hipMalloc((void **)&m.params, sizeof(float) * WSPH_MLP_NPARAM))) return e;
  if ((e = hipMalloc((void **)&m.w1, sizeof(__bf16) * WSPH_MLP_H * WSPH_MLP_IN))) return e;
  if ((e = hipMalloc((void **)&m.w2, sizeof(__bf16) * WSPH_MLP_H * WSPH_MLP_H))) return e;
  if ((e = hipMalloc((void **)&m.w3, sizeof(__bf16) * WSPH_MLP_H))) return e;
  m.nparts = 256;
  if ((e = hipMalloc((void **)&m.part, sizeof(float) * WSPH_MLP_NPARAM * m.nparts))) return e;
  if ((e = hipMalloc((void **)&m.sse_part, sizeof(float) * m.nparts))) return e;
  return hipSuccess;
}

void mlp_free(MlpDev &m) {
  for (void *p : {(void *)m.params, (void *)m.w1, (void *)m.w2, (void *)m.w3, (void *)m.part,
                  (void *)m.sse_part})
    if (p) (void)hipFree(p);
  memset(&m, 0, sizeof(m));
}

void mlp_sync_weights(const MlpDev &m, hipStream_t s) {
  k_to_bf16<<<16, 256, 0, s>>>(m.params, m.w1, m.w2, m.w3);
}

void mlp_forward(const MlpDev &m, const float *x, int n, float *y, hipStream_t s) {
  if (n <= 0) return;
  const int tiles = (n + 63) / 64;
  int blocks = (tiles + kMlpWaves - 1) / kMlpWaves;
  if (blocks > 4096) blocks = 4096;
  k_mlp_forward<<<blocks, 64 * kMlpWaves, 0, s>>>(weights_of(m), x, n, y);
}

void mlp_train_step(MlpDev &m, float *adam_m, float *adam_v, int step, float lr, const float *x,
                    const float *t, int n, float *loss, float *grad, hipStream_t s) {
  const int tiles = (n + 31) / 32;
  int blocks = (tiles + kBwdWaves - 1) / kBwdWaves;
  if (blocks > m.nparts) blocks = m.nparts;
  const float inv_b = 1.0f / (float)n;
  k_mlp_backward<<<blocks, 64 * kBwdWaves, 0, s>>>(weights_of(m), x, t, n, inv_b, m.part,
                                                   m.sse_part);
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  k_adam<<<(WSPH_MLP_NPARAM + 255) / 256, 256, 0, s>>>(m.params, adam_m, adam_v, m.part, blocks,
                                                        grad, m.sse_part, loss, inv_b, lr, b1, b2,
                                                        1e-8f, bc1, bc2);
  if (lr > 0.0f) mlp_sync_weights(m, s);
}

}  // namespace wsph
