// common.h -- device-side helpers shared by the libwsph kernels (gfx950).
//
// All kernels are compiled with -ffp-contract=off: every float32 multiply and
// add rounds separately, in the operation order of the reference kernels
// (wp_sph/solver/IISPH.py, sph_funcs.py), so the results reproduce the CPU
// restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wsph.h"

namespace wsph {

constexpr int kMaxIters = 128;      // upper bound of max_iters / force_iters
constexpr int kHistory = 8192;      // per-step statistics kept on the device (ring)

// Statistics of one step (written by the device, read back on request).
struct StepRecord {
  double sim_time;   // after the step
  double residual;   // mean density error at the end of the Jacobi loop
  double kinetic;    // 0.5 m sum |v|^2 after the step
  double mean_y;     // mean height of the fluid
  float max_y;       // free-surface height: highest fluid particle
  float v_max;       // max |v|
  int iterations;    // Jacobi iterations
  int converged;     // residual <= eta
  int start_source;  // WSPH_WARM_* start value used
  double err_net;    // RMS (learned start - converged p), learned mode
  double err_carry;  // RMS (carried start - converged p), learned mode
};
constexpr int kMaxLdsTable = 16384; // floats per table kept in LDS (64 KiB)
// Learned start: self features and targets are clamped to +-kFeatClamp; a
// particle with |a_ii / a_ref| < kWeakCoupling keeps the carried pressure
// (its pressure barely moves the density and its value is unreliable).
constexpr float kFeatClamp = 64.0f;
// While the network is switched off it is evaluated (and compared) every
// kRetryEvery-th step only.
constexpr int kRetryEvery = 8;
constexpr float kWeakCoupling = 0.05f;
__host__ __device__ inline float clamp_feat(float v) {
  return v > kFeatClamp ? kFeatClamp : (v < -kFeatClamp ? -kFeatClamp : v);
}

// Batched solves: each scene has its own fluid and boundary hash grids,
// stored one after another in the cell tables (slot = base + local slot).
constexpr int kMaxScenes = 4096;
// particles per solver (fluid, and boundary, each): gathers address element
// j of a 16-byte array with the 32-bit byte offset j * 16
constexpr int kMaxParticles = (1 << 28) - 1;
struct SceneGrid {
  int fdims[3], bdims[3];
  int fbase, bbase;
};

// Slots of the list pass's largest neighbour counts (one 128-byte line
// each): every wave of the pass records its maxima, and device-scope atomics
// on one word serialise at the memory side, so the waves spread them.
constexpr int kListMaxSlots = 256;
constexpr int kListMaxStride = 32;  // ints per slot

// status bits accumulated on the device during a step
enum : int {
  ST_FF_OVERFLOW = 1,
  ST_FS_OVERFLOW = 2,
  ST_NAN = 4,
  ST_BARRIER = 8,
};

// Constants of one solver, passed by value to every kernel (the float32
// literals the reference's kernels see, wp_sph/main.py:25-61).
struct SphConst {
  float h;        // SMOOTHING_LENGTH
  float r2_lt;    // sqrtf(d2) < h  <=>  d2 < r2_lt (sqrt is monotone, rounded to nearest)
  float cell_inv; // 1.0f / h, hash grid cell width inverse
  float w_scale;  // float(W_SIZE - 1) * h_inv   (sph_funcs.py:98)
  float g_scale;  // float(GRAD_W_SIZE - 1) * h_inv (sph_funcs.py:113)
  int nw, ng;     // table sizes
  float mass, rho0, omega, inv_small, g_y, vis_mu, xsph;
  int fdims[3], bdims[3]; // hash dims of the fluid and boundary grids
  int layout;             // WSPH_LAYOUT_*
  // feature scales of the pressure predictor: a_ii / a_ref = a_ii * feat_k1 / dt^2,
  // pressures in units of pref
  float feat_k1, pref, inv_pref;
};

// Device-resident step state: clock, residual accumulators, counters.
struct DevState {
  double dt;                  // dt of the coming step
  double sim_time;            // simulated time
  double residual;            // mean density error after the last iteration
  unsigned vmax_bits;         // max |v| (float bits, non-negative)
  int iterations;             // Jacobi iterations of the last solve
  int ap_evals;
  int converged;
  int max_ff, max_fs;         // largest neighbour counts seen
  int status;                 // ST_* bits
  int done;                   // Jacobi loop finished flag
  int iter;                   // current Jacobi iteration (device-driven loop)
  int final_src;              // dp buffer holding the final iterate
  int step_tag;               // incremented by every step's state reset
  int done_tag;               // step_tag of the step whose Jacobi loop finished last
  unsigned bar_count, bar_gen;  // grid barrier
  unsigned bar_seen;            // arrivals seen by a barrier wait that timed out
  int step_count;             // steps recorded in the history ring
  int net_off;                // learned start: 1 = the network lost the last comparison
  int net_eval;               // learned start: the network was evaluated this step
  unsigned long long res_sum[kMaxIters];
  unsigned res_cnt[kMaxIters];
};

// ---- warp.HashGrid cell arithmetic (restated; used at IISPH.py:771-780) ----
__device__ __forceinline__ int hash_axis(int c, int dim) {
  c += 1 << 20;
  c = c < 0 ? 0 : c;
  return c % dim;
}

__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  auto sp = [](uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  return sp(x) | (sp(y) << 1) | (sp(z) << 2);
}

// Slot of hashed cell (hx, hy, hz) in the cell tables.
__device__ __forceinline__ int cell_slot(int hx, int hy, int hz, const int *dims, int layout) {
  if (layout == WSPH_LAYOUT_MORTON) return (int)morton3(hx, hy, hz);
  return hz * (dims[0] * dims[1]) + hy * dims[0] + hx;
}

// Reference linear key cz*dx*dy + cy*dx + cx (hash_grid_index).
__device__ __forceinline__ int linear_key(float x, float y, float z, float inv, const int *dims) {
  const int hx = hash_axis((int)(x * inv), dims[0]);
  const int hy = hash_axis((int)(y * inv), dims[1]);
  const int hz = hash_axis((int)(z * inv), dims[2]);
  return hz * (dims[0] * dims[1]) + hy * dims[0] + hx;
}

// Query range of hash_grid_query on one axis: [trunc((p-r)/w),
// min(trunc((p+r)/w), start + dim - 1)].
struct Range3 {
  int x0, x1, y0, y1, z0, z1;
};
__device__ __forceinline__ Range3 query_range(float px, float py, float pz, float r, float inv,
                                              const int *dims) {
  Range3 q;
  q.x0 = (int)((px - r) * inv);
  q.y0 = (int)((py - r) * inv);
  q.z0 = (int)((pz - r) * inv);
  q.x1 = min((int)((px + r) * inv), q.x0 + dims[0] - 1);
  q.y1 = min((int)((py + r) * inv), q.y0 + dims[1] - 1);
  q.z1 = min((int)((pz + r) * inv), q.z0 + dims[2] - 1);
  return q;
}

// ---- correctly rounded square root and reciprocal -------------------------
// sqrtf(x) for x >= 0 (including +inf): the steps of the compiler's correctly
// rounded expansion (inputs below 2^-96 scaled by 2^32, hardware estimate,
// one-ulp correction each way from the fma residuals, result scaled back by
// 2^-16) without its zero/inf/NaN class test, which these inputs do not need:
// for 0 and +inf both residuals are NaN or zero and the estimate is kept.
// Bit-identical to sqrtf on every non-negative float32 (checked exhaustively
// on gfx950).
__device__ __forceinline__ float sqrt_rn(float x) {
  const bool small = x < 0x1.0p-96f;
  const float xs = small ? x * 0x1.0p+32f : x;
  float s = __builtin_amdgcn_sqrtf(xs);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float ed = __builtin_fmaf(-sd, s, xs);
  const float eu = __builtin_fmaf(-su, s, xs);
  s = ed <= 0.0f ? sd : s;
  s = eu > 0.0f ? su : s;
  return small ? s * 0x1.0p-16f : s;
}

// 1.0f / b rounded to nearest: the hardware reciprocal estimate and one
// Newton step with fma, e = 1 - b*y0 (exact), y = y0 + e*y0. Equal to the
// correctly rounded division for every float32 b in [2^-126, 2^126) (checked
// exhaustively on gfx950); every other b (negative, zero, a reciprocal
// outside the normal range, inf, NaN) takes the division. One unsigned compare
// of the bit pattern selects [2^-126, 2^126) = [0x00800000, 0x7e800000).
__device__ __forceinline__ float rcp_rn(float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  const float e = __builtin_fmaf(-b, y0, 1.0f);
  float y = __builtin_fmaf(e, y0, y0);
  if (__builtin_expect(__float_as_uint(b) - 0x00800000u >= 0x7e000000u, 0)) y = 1.0f / b;
  return y;
}

// ---- kernel tables: spline_W / grad_spline_W (sph_funcs.py:92-118) --------
// The table read is done for every r (the index is clamped to the table) and
// the cut-off applied by a select, so a neighbour loop has no branch.
__device__ __forceinline__ float spline_w(const float *T, int n, float scale, float h, float r) {
  const float idx = r * scale;
  int lo = (int)idx;
  lo = lo < n - 2 ? lo : n - 2;
  const float a = idx - (float)lo;
  const float w = (1.0f - a) * T[lo] + a * T[lo + 1];
  return r >= h ? 0.0f : w;
}

// The square root of grad_w_pair: the steps of sqrt_rn without the range
// scaling, so it equals sqrtf for x >= 2^-96; below that it stays under 2^-47,
// beneath grad_spline_W's 1e-10 cut-off like the exact root (both checked
// exhaustively on gfx950), and the gradient only compares it with the cut-off.
__device__ __forceinline__ float sqrt_rn_pair(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float ed = __builtin_fmaf(-sd, s, x);
  const float eu = __builtin_fmaf(-su, s, x);
  s = ed <= 0.0f ? sd : s;
  return eu > 0.0f ? su : s;
}

// grad of W at x_j - x_i = r, i.e. the table value (-W'/|r|) times r, for a
// pair from the neighbour lists: |r| < h holds by construction
// (the lists test the same float32 radicand with sqrtf(d2) < h), so only the
// |r| < 1e-10 cut-off remains. The cut-off zeroes the table value, giving
// (+-0, +-0, +-0) instead of (+0, +0, +0); every use adds or subtracts the
// gradient (times other factors) into sums that start at +0, where a zero of
// either sign leaves the bits of the sum unchanged.
// grad_w_pair given len, a square root of |r|^2 that is exact from 2^-96 up and
// below 1e-10 under it (sqrt_rn or sqrt_rn_pair).
__device__ __forceinline__ float3 grad_w_len(const float *G, int n, float scale, float len, float3 r) {
  const float idx = len * scale;
  int lo = (int)idx;
  lo = lo < n - 2 ? lo : n - 2;
  const float a = idx - (float)lo;
  float s = (1.0f - a) * G[lo] + a * G[lo + 1];
  s = len < 1e-10f ? 0.0f : s;
  return make_float3(s * r.x, s * r.y, s * r.z);
}
__device__ __forceinline__ float3 grad_w_pair(const float *G, int n, float scale, float3 r) {
  return grad_w_len(G, n, scale, sqrt_rn_pair(r.x * r.x + r.y * r.y + r.z * r.z), r);
}

__device__ __forceinline__ float dist3(float ax, float ay, float az, float bx, float by, float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt_rn(dx * dx + dy * dy + dz * dz);
}

// Element j of a device array through a 32-bit byte offset from its base
// (particle arrays hold at most 2^28 elements, so j * 16 fits), which lets a
// gather use the scalar-base plus 32-bit-offset address form instead of
// 64-bit address arithmetic per lane.
template <typename T>
__device__ __forceinline__ T ld_at(const T *__restrict__ base, int j) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      (unsigned)j * (unsigned)sizeof(T));
}

__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Copy a table into LDS (when small enough) or use it from global memory.
__device__ __forceinline__ const float *stage_table(float *lds, const float *g, int n, bool use_lds) {
  if (!use_lds) return g;
  for (int k = threadIdx.x; k < n; k += blockDim.x) lds[k] = g[k];
  __syncthreads();
  return lds;
}

// a / b rounded to nearest, given y = 1.0f / b rounded to nearest: one
// Markstein correction step, q0 = a*y, r = a - q0*b (exact with fma),
// q = q0 + r*y. Equals IEEE a / b for normal operands (checked against the
// hardware division for every float32 significand of b); used where one
// divisor serves many quotients.
__device__ __forceinline__ float div_by(float a, float b, float y) {
  const float q0 = a * y;
  const float r = __builtin_fmaf(-q0, b, a);
  return __builtin_fmaf(r, y, q0);
}

// Workgroups are dispatched round-robin over the 8 XCDs of the device, each
// with its own L2. Map the hardware block index to a logical tile so that
// every XCD works on one contiguous range of (cell-sorted) particles and the
// neighbour data it gathers stays in its own L2.
#ifndef WSPH_XCD_REMAP
#define WSPH_XCD_REMAP 1
#endif
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile() {
  if (!WSPH_XCD_REMAP) return blockIdx.x;
  const int G = gridDim.x, b = blockIdx.x;
  const int q = G / kXcds, r = G % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
__device__ __forceinline__ int particle_index() { return xcd_tile() * blockDim.x + threadIdx.x; }

// Neighbour lists are tiled by wave: the entries of the 64 particles
// 64w .. 64w + 63 form one block of cap rows of 64 entries, row k holding
// their k-th neighbours. A wave reads row k in one coalesced 256-byte access,
// and its successive rows are adjacent in memory (a column-major list[k * n
// + i] puts them n * 4 bytes apart, each row a fresh DRAM page).
constexpr int kListLanes = 64;
__host__ __device__ inline long list_base(int i, int cap) {
  return ((long)(i / kListLanes) * cap) * kListLanes + (i % kListLanes);
}
// entries of a list of n particles with cap slots (n rounded up to whole blocks)
__host__ __device__ inline size_t list_entries(int n, int cap) {
  return (size_t)((n + kListLanes - 1) / kListLanes) * kListLanes * (size_t)cap;
}

// The fluid and boundary position arrays are allocated with kPointPad entries
// past their last particle: the list builder loads candidate batches ahead of
// the span end, and masks what lies beyond it.
constexpr int kPointPad = 8;

// Visit the neighbours of particle i in list order: entry k at
// list[list_base(i, cap) + k * 64]. The neighbours go in batches
// of kBatch: the batch's gathers are issued together, then the list entries of
// the next batch, and only then is the batch evaluated, so the list load of
// one batch overlaps the gathers of the other. The last, partial batch repeats
// entry cnt - 1 in its unused places (a valid index) and skips them in the
// body. `load(j, slot)` returns the data of neighbour j (slot = its list
// position), `body(data)` accumulates in list order.
template <int kBatch, typename LoadF, typename BodyF>
__device__ __forceinline__ void for_neighbors(const int *__restrict__ list, int cap, int i, int cnt,
                                              LoadF load, BodyF body) {
  if (cnt <= 0) return;
  const int *__restrict__ col = list + list_base(i, cap);
  int jn[kBatch];
#pragma unroll
  for (int u = 0; u < kBatch; ++u) jn[u] = col[min(u, cnt - 1) * kListLanes];
  int k = 0;
  for (; k + kBatch <= cnt; k += kBatch) {
    decltype(load(0, 0)) d[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) d[u] = load(jn[u], k + u);
    if (k + kBatch < cnt) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) jn[u] = col[min(k + kBatch + u, cnt - 1) * kListLanes];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) body(d[u]);
  }
  if (k < cnt) {
    decltype(load(0, 0)) d[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) d[u] = load(jn[u], k + u);
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (k + u < cnt) body(d[u]);
  }
}

// Fixed-point residual term of one particle (WSPH_RESIDUAL_* in wsph.h).
__device__ __forceinline__ unsigned long long residual_fixed(float term) {
  if (term > WSPH_RESIDUAL_TERM_CLAMP) term = WSPH_RESIDUAL_TERM_CLAMP;
  return (unsigned long long)((double)term * WSPH_RESIDUAL_FIXED_SCALE);
}

// wave64 reductions
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
}  // namespace wsph
