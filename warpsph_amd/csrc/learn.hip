// learn.hip -- feature pass of the learned pressure start, with the
// predictor network evaluated inline on the matrix cores.
//
// The reference starts every Jacobi solve from p = 0 (IISPH.py:166-174).
// Here, after k_assemble has written each particle's self features
// (seed = density error, a_ii / a_ref, neighbour count, carried pressure),
// one pass per particle
//   1. gathers the kernel-weighted mean of the four self features over the
//      particle and its fluid neighbours (sum_j W_ij f_j / sum_j W_ij, the
//      particle itself entering with W(0)),
//   2. optionally writes the 8-feature row (training pairs),
//   3. runs the network on the wave's 64 rows in registers (infer64: two
//      32-row MFMA tiles, BF16 operands, FP32 accumulation),
// A network switched off by the arbiter is evaluated every kRetryEvery-th step
// only (DevState::net_eval records whether it ran).
//   4. writes the start value p0 = max(0, carry + P_ref * y) (the carried
//      value where |a_ii / a_ref| < kWeakCoupling, zero inside the a_ii guard
//      band) to aux.z (always) and to the Jacobi iterate dp[0].w
//      unless the network lost the previous step's comparison with the
//      carried pressure (DevState::net_off), in which case dp[0].w keeps the
//      carried value k_assemble wrote.
#include "kernels.h"
#include "mlp_dev.h"

namespace wsph {

constexpr int kLearnWaves = 4;  // one block = 256 particles = 8 MFMA tiles

struct PosSeed {
  float4 pos, seed;
};

constexpr int kLearnTable = 4096;  // W table entries kept in LDS (larger tables: global)

template <bool kInfer>
__global__ __launch_bounds__(64 * kLearnWaves, 4) void k_learned_start(Bufs b, SphConst c, Weights W) {
  __shared__ NetLds L;
  __shared__ float wt_lds[kLearnTable];
  bool eval = false;
  if constexpr (kInfer) {
    // a switched-off network is retried every kRetryEvery-th step (the same
    // decision in every block); without logging nothing else is needed
    eval = !b.st->net_off || b.st->step_count % kRetryEvery == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) b.st->net_eval = eval ? 1 : 0;
    if (!eval && !b.feat) return;
    if (eval) stage_net(L, W);
  }
  // the kernel table in LDS: each lane's lookups hit other addresses
  const bool tab_lds = c.nw <= kLearnTable;
  if (tab_lds)
    for (int k = threadIdx.x; k < c.nw; k += blockDim.x) wt_lds[k] = b.wt[k];
  __syncthreads();
  const float *wtab = tab_lds ? wt_lds : b.wt;
  const int i = particle_index();
  const bool live = i < b.n;  // whole waves stay to the end: infer64 exchanges lanes
  float f[WSPH_MLP_IN] = {};
  float carry = 0.0f, a_ii = 0.0f;
  if (live) {
    const int n = b.n;
    const float4 p = b.pos[i];
    const float4 s = b.seed[i];
    const float w0 = spline_w(wtab, c.nw, c.w_scale, c.h, 0.0f);
    float sw = w0, m0 = w0 * s.x, m1 = w0 * s.y, m2 = w0 * s.z, m3 = w0 * s.w;
    const float4 *__restrict__ pos = b.pos;
    const float4 *__restrict__ seed = b.seed;
    for_neighbors<4>(
        b.nbr_ff, b.cap_ff, i, b.cnt_ff[i],
        [&](int j, int) { return PosSeed{ld_at(pos, j), ld_at(seed, j)}; },
        [&](const PosSeed &v) {
          const float w = spline_w(wtab, c.nw, c.w_scale, c.h,
                                   dist3(p.x, p.y, p.z, v.pos.x, v.pos.y, v.pos.z));
          sw += w;
          m0 += w * v.seed.x;
          m1 += w * v.seed.y;
          m2 += w * v.seed.z;
          m3 += w * v.seed.w;
        });
    const float inv = 1.0f / sw;
    f[0] = s.x;
    f[1] = s.y;
    f[2] = s.z;
    f[3] = s.w;
    f[4] = m0 * inv;
    f[5] = m1 * inv;
    f[6] = m2 * inv;
    f[7] = m3 * inv;
    carry = b.pprev[i];
    a_ii = b.aux[i].y;
    if (b.feat) {
      float4 *row = reinterpret_cast<float4 *>(b.feat + (long)i * WSPH_MLP_IN);
      row[0] = make_float4(f[0], f[1], f[2], f[3]);
      row[1] = make_float4(f[4], f[5], f[6], f[7]);
    }
  }
  if constexpr (kInfer) {
    if (!eval) return;
    const int net_off = b.st->net_off;
    const float y = infer64(L, f);
    if (live) {
      float pn = isfinite(y) ? carry + c.pref * y : carry;
      pn = pn > 0.0f ? pn : 0.0f;
      if (fabsf(f[1]) < kWeakCoupling) pn = carry;
      if (a_ii > -c.inv_small && a_ii < c.inv_small) pn = 0.0f;
      reinterpret_cast<float *>(&b.aux[i])[2] = pn;
      if (!net_off) reinterpret_cast<float *>(&b.dp[0][i])[3] = pn;
    }
  }
}

void launch_learned_start(const Bufs &b, const SphConst &c, const Weights *net, hipStream_t s) {
  if (b.n == 0) return;
  const int block = 64 * kLearnWaves;
  const int grid = (b.n + block - 1) / block;
  if (net)
    k_learned_start<true><<<grid, block, 0, s>>>(b, c, *net);
  else
    k_learned_start<false><<<grid, block, 0, s>>>(b, c, Weights{});
}

}  // namespace wsph
