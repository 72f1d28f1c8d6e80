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

// SiLU and its derivative with the hardware reciprocal (one ulp; the operands
// of every layer are rounded to bf16 anyway): the library is built with
// correctly rounded division, which would cost a full division sequence per
// activation.
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float dsilu(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
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

// ---- register-resident inference of 64 rows per wave --------------------
// The transposed products z^T = W x^T are computed (M = hidden units,
// N = the wave's rows), so that a layer's accumulator, in D layout, becomes
// the next layer's B operand with one exchange of four values between lanes
// l and l ^ 32 per 16-wide k-slice: no LDS round trip and no block barrier.
// The weights sit in LDS (staged once per block), each lane reading its A
// operand rows as 16-byte loads; the output layer is a dot product on the
// VALU (bf16 operands, exact products, FP32 sums, like the matrix cores).
constexpr int kW2Pitch = WSPH_MLP_H + 8;  // bf16 row pitch of W2 in LDS (breaks bank conflicts)
struct NetLds {
  __bf16 w1[WSPH_MLP_H * WSPH_MLP_IN];
  __bf16 w2[WSPH_MLP_H * kW2Pitch];
  float b1[WSPH_MLP_H], b2[WSPH_MLP_H], w3[WSPH_MLP_H];  // w3 holds the bf16 values
  float b3;
};

// Every thread of the block calls this once; a barrier follows.
__device__ __forceinline__ void stage_net(NetLds &L, const Weights &W) {
  for (int k = threadIdx.x; k < WSPH_MLP_H * WSPH_MLP_H; k += blockDim.x) {
    L.w2[(k / WSPH_MLP_H) * kW2Pitch + k % WSPH_MLP_H] = W.w2[k];
    if (k < WSPH_MLP_H * WSPH_MLP_IN) L.w1[k] = W.w1[k];
    if (k < WSPH_MLP_H) {
      L.b1[k] = W.b1[k];
      L.b2[k] = W.b2[k];
      L.w3[k] = (float)W.w3[k];
    }
  }
  if (threadIdx.x == 0) L.b3 = W.b3[0];
}

// D-layout bias of M-tile t: register i of lane (n, h) holds unit
// m = 32 t + 8 (i / 4) + 4 h + i % 4.
__device__ __forceinline__ f32x16 bias_d(const float *b, int t, int h) {
  f32x16 c;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 v = *reinterpret_cast<const float4 *>(b + 32 * t + 8 * g + 4 * h);
    c[4 * g] = v.x;
    c[4 * g + 1] = v.y;
    c[4 * g + 2] = v.z;
    c[4 * g + 3] = v.w;
  }
  return c;
}

// y for the row of every lane of the wave (lane l holds row l's 8 features
// in f). Rows 0-31 form the first tile, rows 32-63 the second.
__device__ __forceinline__ float infer64(const NetLds &L, const float f[WSPH_MLP_IN]) {
  const int lane = threadIdx.x & 63, h = lane >> 5, r = lane & 31;
  float yt[2];
  // one tile, and within it one accumulator, live at a time (register budget)
#pragma unroll 1
  for (int tile = 0; tile < 2; ++tile) {
    // an offset the compiler cannot see through: the weight loads stay in
    // this iteration instead of being hoisted (and kept live) across tiles
    int salt = 0;
    asm volatile("" : "+v"(salt));
    const __bf16 *w1 = L.w1 + salt, *w2 = L.w2 + salt;
    const float *b1 = L.b1 + salt, *b2 = L.b2 + salt, *w3 = L.w3 + salt;
    // B of layer 1: x^T[k][n], lanes 0-31 hold row n's features, 32-63 the
    // zero padding of K = 8 to 16
    bf16x8 bx;
#pragma unroll
    for (int e = 0; e < WSPH_MLP_IN; ++e) {
      const float v = tile == 0 ? f[e] : __shfl_xor(f[e], 32, 64);
      bx[e] = h == 0 ? (__bf16)v : (__bf16)0.0f;
    }
    // layer 1, M-tile by M-tile; its activations (D layout) become the B
    // operands of layer 2's k-slices 2t and 2t + 1
    bf16x8 bh[4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bf16x8 aw;
      if (h == 0) {
        aw = *reinterpret_cast<const bf16x8 *>(&w1[(32 * t + r) * WSPH_MLP_IN]);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) aw[e] = (__bf16)0.0f;
      }
      const f32x16 d = mfma(aw, bx, bias_d(b1, t, h));
      float a[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] = silu(d[i]);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int g0 = 2 * q;
        float recv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) recv[e] = __shfl_xor(h == 0 ? a[4 * (g0 + 1) + e] : a[4 * g0 + e], 32, 64);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          bh[2 * t + q][e] = (__bf16)(h == 0 ? a[4 * g0 + e] : recv[e]);
          bh[2 * t + q][e + 4] = (__bf16)(h == 0 ? recv[e] : a[4 * (g0 + 1) + e]);
        }
      }
    }
    // layer 2, M-tile by M-tile, folded into the output layer's dot product
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      f32x16 d = bias_d(b2, t, h);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 aw = *reinterpret_cast<const bf16x8 *>(
            &w2[(32 * t + r) * kW2Pitch + 16 * ks + 8 * h]);
        d = mfma(aw, bh[ks], d);
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 w = *reinterpret_cast<const float4 *>(w3 + 32 * t + 8 * g + 4 * h);
        part += (float)(__bf16)silu(d[4 * g]) * w.x;
        part += (float)(__bf16)silu(d[4 * g + 1]) * w.y;
        part += (float)(__bf16)silu(d[4 * g + 2]) * w.z;
        part += (float)(__bf16)silu(d[4 * g + 3]) * w.w;
      }
    }
    yt[tile] = part + __shfl_xor(part, 32, 64) + L.b3;
  }
  return h == 0 ? yt[0] : yt[1];
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
