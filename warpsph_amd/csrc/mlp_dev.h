// mlp_dev.h -- device building blocks of the pressure predictor on the gfx950
// matrix cores (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation).
//
// One wave evaluates a tile of 32 rows. Operand layout of the 32x32x16 MFMA
// (checked on the device): A[m][k] sits in lane m + 32*(k/8), element k%8;
// B[k][n] in lane n + 32*(k/8), element k%8; D[m][n] in lane
// n + 32*((m/4)%2), register 4*(m/8) + m%4. Activations move between layers
// through LDS, where they are re-read in the A (or B) layout. forward_tile
// contains block barriers: every wave of the block must call it the same
// number of times.
#pragma once

#include <hip/hip_runtime.h>

#include "mlp.h"

namespace wsph {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// row of accumulator register i for this lane
__device__ __forceinline__ int drow(int lane, int i) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Load the A operand (rows 0..31, k-slice ks of 16) from a float matrix
// m[32][ld] in LDS.
__device__ __forceinline__ bf16x8 lds_a(const float (*m)[65], int lane, int ks) {
  bf16x8 v;
  const int r = lane & 31, k0 = 16 * ks + 8 * (lane >> 5);
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)m[r][k0 + j];
  return v;
}

// B operand of W^T for output tile nt (columns 32*nt..) and k-slice ks,
// W stored row-major [nout][nin] in bf16 (global or LDS).
__device__ __forceinline__ bf16x8 w_b(const __bf16 *w, int nin, int nout, int lane, int nt, int ks) {
  bf16x8 v;
  const int n = 32 * nt + (lane & 31), k0 = 16 * ks + 8 * (lane >> 5);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = k0 + j;
    v[j] = (n < nout && k < nin) ? w[n * nin + k] : (__bf16)0.0f;
  }
  return v;
}

// Forward pass of one 32-row tile. x: features (fp32, row-major, 8 per row;
// rows past `rows` are read as zero). Leaves h2 in act ([32][65] floats of LDS); when z1s / z2s are
// given, the pre-activations of layers 1 and 2 are stored there ([32][64]).
// Returns y in y[16]: lane 0 holds rows drow(0,i), lane 32 rows drow(32,i).
struct Weights {
  const __bf16 *w1, *w2, *w3;  // bf16 copies [64][8], [64][64], [64]
  const float *b1, *b2, *b3;   // fp32 biases
};

__device__ __forceinline__ void forward_tile(const Weights &W, const float *x, int rows,
                                             float (*act)[65], float *z1s, float *z2s,
                                             float y[16]) {
  const int lane = threadIdx.x & 63;
  // layer 1: A = x (K = 8, padded to 16)
  bf16x8 ax;
  {
    const int r = lane & 31, k0 = 8 * (lane >> 5);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + j;
      ax[j] = (r < rows && k < WSPH_MLP_IN) ? (__bf16)x[r * WSPH_MLP_IN + k] : (__bf16)0.0f;
    }
  }
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    f32x16 c;
    const int n = 32 * nt + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = W.b1[n];
    c = mfma(ax, w_b(W.w1, WSPH_MLP_IN, WSPH_MLP_H, lane, nt, 0), c);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = drow(lane, i);
      act[r][n] = silu(c[i]);
      if (z1s) z1s[r * 64 + n] = c[i];
    }
  }
  __syncthreads();
  // layer 2
  bf16x8 ah[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) ah[ks] = lds_a(act, lane, ks);
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    f32x16 c;
    const int n = 32 * nt + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = W.b2[n];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) c = mfma(ah[ks], w_b(W.w2, WSPH_MLP_H, WSPH_MLP_H, lane, nt, ks), c);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = drow(lane, i);
      act[r][n] = silu(c[i]);
      if (z2s) z2s[r * 64 + n] = c[i];
    }
  }
  __syncthreads();
  // layer 3: y = h2 . w3 + b3 (one output column: B has w3 in column 0)
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) ah[ks] = lds_a(act, lane, ks);
  f32x16 c;
#pragma unroll
  for (int i = 0; i < 16; ++i) c[i] = W.b3[0];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) c = mfma(ah[ks], w_b(W.w3, WSPH_MLP_H, 1, lane, 0, ks), c);
#pragma unroll
  for (int i = 0; i < 16; ++i) y[i] = c[i];
  __syncthreads();
}

inline Weights weights_of(const MlpDev &m) {
  Weights W;
  W.w1 = m.w1;
  W.w2 = m.w2;
  W.w3 = m.w3;
  W.b1 = m.params + WSPH_MLP_OFF_B1;
  W.b2 = m.params + WSPH_MLP_OFF_B2;
  W.b3 = m.params + WSPH_MLP_OFF_B3;
  return W;
}

}  // namespace wsph
