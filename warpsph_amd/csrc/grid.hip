// grid.hip -- uniform hash grid build on the device (gfx950).
//
// Replaces warp.HashGrid.build (called at IISPH.py:811 and 1414; warp-lang
// 1.7.1 internals restated in common.h): every particle gets the cell
// trunc(p / h) hashed exactly like hash_grid_index, particles are counting-
// sorted by cell slot (wave64 prefix scan over the cell histogram), and
// inside each cell they are ordered by original index, which is the order
// the reference's stable radix sort yields. The particle arrays themselves
// are then permuted into cell order so that neighbour gathers stay local.
#include "kernels.h"

namespace wsph {

__global__ void k_fill_u32(unsigned *p, unsigned v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Cell slot of every particle plus its arrival rank in that cell.
// With scenes, the point's own scene grid is used and its slot offset by the
// scene's base, so the scenes' cells never mix. The particles arrive in last
// step's cell order, so a wave holds a few runs of equal slots: one atomic per
// distinct slot and wave hands out the ranks (their order inside a cell does
// not matter: k_order_in_cell orders each cell by original index).
__global__ void k_cell_slots(const float4 *__restrict__ pos, int n, float inv, int d0, int d1,
                             int d2, int layout, int *__restrict__ slot, int *__restrict__ rank,
                             unsigned *__restrict__ count, const int *__restrict__ sid,
                             const SceneGrid *__restrict__ scenes, int boundary) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const float4 p = live ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int dims[3] = {d0, d1, d2};
  int base = 0;
  if (scenes && live) {
    const SceneGrid &g = scenes[sid[i]];
    const int *d = boundary ? g.bdims : g.fdims;
    dims[0] = d[0];
    dims[1] = d[1];
    dims[2] = d[2];
    base = boundary ? g.bbase : g.fbase;
  }
  const int hx = hash_axis((int)(p.x * inv), dims[0]);
  const int hy = hash_axis((int)(p.y * inv), dims[1]);
  const int hz = hash_axis((int)(p.z * inv), dims[2]);
  const int c = live ? base + cell_slot(hx, hy, hz, dims, layout) : -1;
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(live);
  int r = 0;
  while (pending) {  // wave-uniform: one pass per distinct slot of the wave
    const int leader = __ffsll((long long)pending) - 1;
    const int cl = __shfl(c, leader, 64);
    const unsigned long long same = __ballot(c == cl);
    unsigned base_rank = 0;
    if (lane == leader) base_rank = atomicAdd(&count[cl], (unsigned)__popcll(same));
    base_rank = __shfl(base_rank, leader, 64);
    if (c == cl) r = (int)base_rank + __popcll(same & ((1ull << lane) - 1ull));
    pending &= ~same;
  }
  if (live) {
    slot[i] = c;
    rank[i] = r;
  }
}

__global__ void k_place(const int *__restrict__ slot, const int *__restrict__ rank,
                        const unsigned *__restrict__ start, int n, int *__restrict__ tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tmp[start[slot[i]] + rank[i]] = i;
}

// Final position inside the cell = number of cell members with a smaller
// original index; perm[final] = source index.
__global__ void k_order_in_cell(const int *__restrict__ tmp, const int *__restrict__ slot,
                                const unsigned *__restrict__ start,
                                const unsigned *__restrict__ count, const int *__restrict__ orig,
                                int n, int *__restrict__ perm) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = tmp[t];
  const int c = slot[i];
  const int s = (int)start[c], e = s + (int)count[c];
  const int oi = orig[i];
  int r = 0;
  for (int u = s; u < e; ++u) r += orig[tmp[u]] < oi;
  perm[s + r] = i;
}

__global__ void k_gather_fluid(const int *__restrict__ perm, int n, const float4 *__restrict__ pos,
                               const float4 *__restrict__ vel, const float *__restrict__ pprev,
                               const int *__restrict__ orig, float4 *__restrict__ pos_o,
                               float4 *__restrict__ vel_o, float *__restrict__ pprev_o,
                               int *__restrict__ orig_o) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int i = perm[f];
  pos_o[f] = pos[i];
  vel_o[f] = vel[i];
  pprev_o[f] = pprev[i];
  orig_o[f] = orig[i];
}

// ---- exclusive prefix scan (wave64 shuffles + LDS across waves) -----------
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanChunk = kScanThreads * kScanPerThread;

__global__ __launch_bounds__(kScanThreads) void k_scan_chunk(const unsigned *__restrict__ in,
                                                             unsigned *__restrict__ out,
                                                             unsigned *__restrict__ chunk_sums,
                                                             long n) {
  __shared__ unsigned wave_tot[kScanThreads / 64];
  const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanPerThread;
  unsigned v[kScanPerThread];
  unsigned t = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    t += v[k];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned x = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_tot[wid] = x;
  __syncthreads();
  unsigned wave_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    const unsigned s = wave_tot[w];
    if (w < wid) wave_off += s;
    total += s;
  }
  unsigned run = wave_off + x - t;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (chunk_sums && threadIdx.x == 0) chunk_sums[blockIdx.x] = total;
}

__global__ void k_scan_add(unsigned *__restrict__ out, const unsigned *__restrict__ offs, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += offs[i / kScanChunk];
}

int scan_levels_needed(long n, long *sizes, int max_levels) {
  int L = 0;
  while (n > kScanChunk && L < max_levels) {
    n = (n + kScanChunk - 1) / kScanChunk;
    sizes[L++] = n;
  }
  return L;
}

// out = exclusive scan of in[0..n); level_bufs[l] hold the chunk sums of
// each recursion level (sizes from scan_levels_needed), two per level.
void exclusive_scan(const unsigned *in, unsigned *out, long n, unsigned **level_bufs,
                    hipStream_t st) {
  const long nchunk = (n + kScanChunk - 1) / kScanChunk;
  if (nchunk <= 1) {
    k_scan_chunk<<<1, kScanThreads, 0, st>>>(in, out, nullptr, n);
    return;
  }
  unsigned *sums = level_bufs[0];
  unsigned *sums_scan = level_bufs[1];
  k_scan_chunk<<<(unsigned)nchunk, kScanThreads, 0, st>>>(in, out, sums, n);
  exclusive_scan(sums, sums_scan, nchunk, level_bufs + 2, st);
  k_scan_add<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(out, sums_scan, n);
}

// Sort particles into cell order. cur arrays are read, the other buffer
// set is written, and cell start/count tables are filled.
void sort_into_cells(const GridArgs &g, hipStream_t st) {
  const int n = g.n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  k_fill_u32<<<(unsigned)((g.ncell + 255) / 256), 256, 0, st>>>(g.count, 0u, g.ncell);
  if (n > 0)
    k_cell_slots<<<nb, 256, 0, st>>>(g.pos, n, g.inv, g.dims[0], g.dims[1], g.dims[2], g.layout,
                                     g.slot, g.rank, g.count, g.sid, g.scenes, g.boundary);
  exclusive_scan(g.count, g.start, g.ncell, g.scan_bufs, st);
  if (n == 0) return;
  k_place<<<nb, 256, 0, st>>>(g.slot, g.rank, g.start, n, g.tmp);
  k_order_in_cell<<<nb, 256, 0, st>>>(g.tmp, g.slot, g.start, g.count, g.orig, n, g.perm);
}

void gather_fluid(const int *perm, int n, const float4 *pos, const float4 *vel, const float *pprev,
                  const int *orig, float4 *pos_o, float4 *vel_o, float *pprev_o, int *orig_o,
                  hipStream_t st) {
  if (n == 0) return;
  k_gather_fluid<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(perm, n, pos, vel, pprev, orig,
                                                               pos_o, vel_o, pprev_o, orig_o);
}

__global__ void k_gather_i32(const int *__restrict__ perm, int n, const int *__restrict__ src,
                             int *__restrict__ dst) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) dst[f] = src[perm[f]];
}

void gather_i32(const int *perm, int n, const int *src, int *dst, hipStream_t st) {
  if (n > 0) k_gather_i32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(perm, n, src, dst);
}

void fill_u32(unsigned *p, unsigned v, long n, hipStream_t st) {
  if (n > 0) k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, v, n);
}

__global__ void k_linear_keys(const float *__restrict__ x, int n, float inv, int d0, int d1,
                              int d2, int *__restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int dims[3] = {d0, d1, d2};
  keys[i] = linear_key(x[3 * (long)i], x[3 * (long)i + 1], x[3 * (long)i + 2], inv, dims);
}

void linear_keys(const float *x, int n, float inv, const int *dims, int *keys, hipStream_t st) {
  if (n > 0)
    k_linear_keys<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x, n, inv, dims[0], dims[1],
                                                                dims[2], keys);
}

}  // namespace wsph
