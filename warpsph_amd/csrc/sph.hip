// sph.hip -- IISPH neighbour passes on gfx950.
//
// One thread per particle (cell-sorted order). Every sum runs over the
// particle's neighbours in the reference's enumeration order and every
// expression keeps the reference's operation order (IISPH.py:23-699), so the
// float32 results equal those of the CPU restatement exactly. Neighbour
// lists are tiled by wave (entry k of particle i at list[list_base(i, cap) +
// k * 64]) so that a wave reads slot k of its 64 particles in one coalesced
// access. The kernel
// tables sit in LDS when they fit (sizes up to kMaxLdsTable floats).
#include "kernels.h"

namespace wsph {

extern __shared__ float g_smem[];

struct Tabs {
  const float *w, *g;
};

template <bool kLds>
__device__ __forceinline__ Tabs stage_tabs(const Bufs &b, const SphConst &c) {
  if constexpr (kLds) {
    for (int k = threadIdx.x; k < c.nw; k += blockDim.x) g_smem[k] = b.wt[k];
    for (int k = threadIdx.x; k < c.ng; k += blockDim.x) g_smem[c.nw + k] = b.gt[k];
    __syncthreads();
    return Tabs{g_smem, g_smem + c.nw};
  } else {
    return Tabs{b.wt, b.gt};
  }
}

// Neighbours loaded ahead per thread, per pass (J1 is ALU-bound, the other
// passes wait on memory).
#ifndef WSPH_BATCH_J1
#define WSPH_BATCH_J1 4
#endif
#ifndef WSPH_BATCH_J2
#define WSPH_BATCH_J2 4
#endif
#ifndef WSPH_BATCH_OTHER
#define WSPH_BATCH_OTHER 4
#endif
constexpr int kBatch = WSPH_BATCH_OTHER;
constexpr int kBatchJ1 = WSPH_BATCH_J1;
constexpr int kBatchJ2 = WSPH_BATCH_J2;

struct F4x2 {
  float4 a, b;
};
#define W_OF(T, r) spline_w((T).w, c.nw, c.w_scale, c.h, (r))
// every GW_OF pair comes from the neighbour lists
#define GW_OF(T, r) grad_w_pair((T).g, c.ng, c.g_scale, (r))

// (dx*dx + dy*dy) + dz*dz with d = a - b: the radicand of wp.length(a - b)
__device__ __forceinline__ float dist2(float ax, float ay, float az, float bx, float by, float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ float3 sub3(float4 a, float4 b) {
  return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}

// update_p, IISPH.py:521-539
__device__ __forceinline__ float update_p(float p, float rho_adv, float ap, float a_ii,
                                          const SphConst &c) {
  float a = a_ii;
  if (a > -c.inv_small && a < c.inv_small) a = (a < 0.0f ? -1.0f : 1.0f) * c.inv_small;
  const float pi = p + (c.rho0 - rho_adv - ap) * c.omega / a;
  return pi > 0.0f ? pi : 0.0f;
}

// ---- count_/store_same_neighbor (IISPH.py:52-104) and
//      count_/store_diff_neighbor (IISPH.py:107-163) --------------------------
// Candidates of one query (hash_grid_query, cells x innermost, then y, then
// z; ascending original index inside a cell) are visited through `visit`.
// With the linear cell layout the x-cells of one (y, z) row occupy
// consecutive slots, so their particles form one contiguous range and a row
// is a single span; otherwise each cell is its own span.
template <typename VisitF>
__device__ __forceinline__ void for_query_spans(float px, float py, float pz, const SphConst &c,
                                                const int *dims, const unsigned *__restrict__ start,
                                                const unsigned *__restrict__ count, VisitF visit) {
  const Range3 q = query_range(px, py, pz, c.h, c.cell_inv, dims);
  for (int z = q.z0; z <= q.z1; ++z) {
    const int hz = hash_axis(z, dims[2]);
    for (int y = q.y0; y <= q.y1; ++y) {
      const int hy = hash_axis(y, dims[1]);
      const int hx0 = hash_axis(q.x0, dims[0]);
      const int hx1 = hash_axis(q.x1, dims[0]);
      if (c.layout == WSPH_LAYOUT_LINEAR && hx1 - hx0 == q.x1 - q.x0) {
        const int s0 = cell_slot(hx0, hy, hz, dims, c.layout);
        const int s1 = s0 + (hx1 - hx0);
        visit((int)start[s0], (int)(start[s1] + count[s1]));
      } else {
        for (int x = q.x0; x <= q.x1; ++x) {
          const int sl = cell_slot(hash_axis(x, dims[0]), hy, hz, dims, c.layout);
          visit((int)start[sl], (int)(start[sl] + count[sl]));
        }
      }
    }
  }
}

// The hash grid a query of particle i searches: the solver's grid, or the
// grid of the particle's own scene in a batch (its cell tables start at the
// scene's base slot).
struct GridRef {
  int dims[3];
  const unsigned *start, *count;
};
__device__ __forceinline__ GridRef fluid_grid(const Bufs &b, const SphConst &c, int scene) {
  GridRef g;
  if (b.scenes) {
    const SceneGrid &sg = b.scenes[scene];
    g.dims[0] = sg.fdims[0];
    g.dims[1] = sg.fdims[1];
    g.dims[2] = sg.fdims[2];
    g.start = b.fstart + sg.fbase;
    g.count = b.fcount + sg.fbase;
  } else {
    g.dims[0] = c.fdims[0];
    g.dims[1] = c.fdims[1];
    g.dims[2] = c.fdims[2];
    g.start = b.fstart;
    g.count = b.fcount;
  }
  return g;
}
__device__ __forceinline__ GridRef boundary_grid(const Bufs &b, const SphConst &c, int scene) {
  GridRef g;
  if (b.scenes) {
    const SceneGrid &sg = b.scenes[scene];
    g.dims[0] = sg.bdims[0];
    g.dims[1] = sg.bdims[1];
    g.dims[2] = sg.bdims[2];
    g.start = b.bstart + sg.bbase;
    g.count = b.bcount + sg.bbase;
  } else {
    g.dims[0] = c.bdims[0];
    g.dims[1] = c.bdims[1];
    g.dims[2] = c.bdims[2];
    g.start = b.bstart;
    g.count = b.bcount;
  }
  return g;
}

// Candidates [s, e) tested kBatch at a time (positions loaded ahead).
template <typename TestF>
__device__ __forceinline__ void for_span(const float4 *__restrict__ pts, int s, int e, TestF test) {
  int j = s;
  for (; j + kBatch <= e; j += kBatch) {
    float4 o[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) o[u] = pts[j + u];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) test(j + u, o[u]);
  }
  for (; j < e; ++j) test(j, pts[j]);
}

// hash_axis with the modulo as a float-reciprocal division: for
// 0 <= c < 2^22 and 1 <= dim <= 1024, q = (int)((float)c * (1.0f / dim)) is
// off by at most one, which one correction in each direction removes (checked
// exhaustively over that whole range); larger c keep the integer modulo.
__device__ __forceinline__ int hash_axis_fast(int c, int dim, float inv) {
  c += 1 << 20;
  c = c < 0 ? 0 : c;
  if (c >= (1 << 22)) return c % dim;
  const int q = (int)((float)c * inv);
  int r = c - q * dim;
  r = r < 0 ? r + dim : r;
  return r >= dim ? r - dim : r;
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// a[k] for k in [0, 4) by selects (no indexed register array, which would
// live in scratch memory inside a loop that is not unrolled)
__device__ __forceinline__ uint32_t pick4(const uint32_t (&a)[4], int k) {
  return k == 0 ? a[0] : k == 1 ? a[1] : k == 2 ? a[2] : a[3];
}

// The traversal of for_query_spans (same cells, same order) with wave-uniform
// loop counts: every lane runs as many (z, y) rows as the widest query range
// of its wave (rows outside a lane's own range visit nothing), and `row_done`
// runs after each row with the whole wave converged. The hashed coordinates
// of the (at most 4) cells per axis and their Morton bit spreads are computed
// once per query. Lanes with live == false visit nothing.
template <typename VisitF, typename RowF>
__device__ __forceinline__ void for_query_rows(float px, float py, float pz, bool live,
                                               const SphConst &c, const int *dims,
                                               const unsigned *__restrict__ start,
                                               const unsigned *__restrict__ count, VisitF visit,
                                               RowF row_done) {
  const Range3 q = query_range(px, py, pz, c.h, c.cell_inv, dims);
  const int d0 = dims[0], d1 = dims[1], d2 = dims[2];
  const float i0 = 1.0f / (float)d0, i1 = 1.0f / (float)d1, i2 = 1.0f / (float)d2;
  const bool morton = c.layout == WSPH_LAYOUT_MORTON;
  uint32_t kx[4], ky[4], kz[4];  // per-axis slot parts of the range's cells
  const int hx0 = hash_axis_fast(q.x0, d0, i0), hx1 = hash_axis_fast(q.x1, d0, i0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int hx = hash_axis_fast(q.x0 + k, d0, i0);
    const int hy = hash_axis_fast(q.y0 + k, d1, i1);
    const int hz = hash_axis_fast(q.z0 + k, d2, i2);
    kx[k] = morton ? spread3((uint32_t)hx) : (uint32_t)hx;
    ky[k] = morton ? spread3((uint32_t)hy) << 1 : (uint32_t)(hy * d0);
    kz[k] = morton ? spread3((uint32_t)hz) << 2 : (uint32_t)(hz * d0 * d1);
  }
  const int nxl = q.x1 - q.x0 + 1;
  const int nz = wave_max_i32(live ? q.z1 - q.z0 + 1 : 0);
  const int ny = wave_max_i32(live ? q.y1 - q.y0 + 1 : 0);
  for (int dz = 0; dz < nz; ++dz) {
    for (int dy = 0; dy < ny; ++dy) {
      if (live && q.z0 + dz <= q.z1 && q.y0 + dy <= q.y1) {
        const uint32_t zy = pick4(kz, dz) + pick4(ky, dy);
        if (!morton && hx1 - hx0 == nxl - 1) {
          // linear layout, no wrap: the row's cells are consecutive slots
          const int s0 = (int)(zy + kx[0]);
          const int s1 = s0 + (nxl - 1);
          visit((int)start[s0], (int)(start[s1] + count[s1]));
        } else {
          for (int dx = 0; dx < nxl; ++dx) {
            const uint32_t x = pick4(kx, dx);
            const int sl = (int)(morton ? (zy | x) : zy + x);
            visit((int)start[sl], (int)(start[sl] + count[sl]));
          }
        }
      }
      row_done();
    }
  }
}

// Neighbour list writer of one wave. Entry k of lane l belongs at
// list[list_base(i_l, cap) + k * 64]; the lanes find their entries at different times, and
// storing each as found fills a 128-byte line of slot k from ~32 separate
// partial stores, between which the line is evicted and written back again
// under the candidate-read traffic. Entries are therefore kept in a ring of
// kListRing slot rows per wave in LDS, and a slot row is stored once, as one
// coalesced row, when every lane of the wave has passed it (after each query
// row) or at the end. A lane more than kListRing slots ahead of the wave's
// flushed rows stores its later entries directly.
constexpr int kListRing = 32;  // a power of two
static_assert((kListRing & (kListRing - 1)) == 0, "ring rows: a power of two");
constexpr unsigned kRingRowMask = (kListRing - 1) * 256u;  // byte offset of a ring row
typedef __attribute__((address_space(3))) int lds_int;  // an int in shared memory
#ifndef WSPH_LIST_BATCH
#define WSPH_LIST_BATCH 4
#endif
constexpr int kListBatch = WSPH_LIST_BATCH;  // candidates loaded together per lane
static_assert(kListBatch <= kPointPad, "a batch reads at most kPointPad places past a span");
struct ListWriter {
  int *ring;      // [kListRing][64] of this wave
  int *list;
  long base;      // list_base(i, cap): entry k at list[base + k * 64]
  int i, lane, cap;
  bool live;
  // Entries found so far times 256: entry k's ring slot lies at byte
  // (k * 256) mod (kListRing * 256) + lane * 4 of the wave's ring, so the
  // running count is also the slot's row offset.
  unsigned cb = 0;
  int direct = 0x7fffffff, flushed = 0;
  // put_batch may store entries up to lim - 1 (ring rows not yet flushed and
  // below the capacity), times 256; 0 once entries go straight to the list.
  unsigned limb = 0;
  // shared-memory address of this lane's slot in ring row 0; put_batch ORs
  // the row offset into it, which needs the wave's ring at a multiple of
  // kListRing * 256 bytes (else put_batch is not used)
  unsigned lane_at = 0;
  __device__ int cnt() const { return (int)(cb >> 8); }
  __device__ void init() {
    const unsigned at = (unsigned)(size_t)(lds_int *)ring;
    lane_at = at + (unsigned)lane * 4u;
    limb = (at & (kListRing * 256u - 1u)) == 0u ? (unsigned)(kListRing < cap ? kListRing : cap) << 8
                                                : 0u;
  }
  __device__ void put(int j) {
    const int k = cnt();
    if (k < cap) {
      if (direct == 0x7fffffff && k - flushed < kListRing) {
        ring[((unsigned)k & (kListRing - 1)) * 64 + lane] = j;
      } else {
        if (direct == 0x7fffffff) direct = k;
        list[base + (long)k * kListLanes] = j;
        limb = 0;
      }
    }
    cb += 256u;
  }
  // Room for k more entries in the ring: then put_batch may be used.
  __device__ bool room(int k) const { return cb + ((unsigned)k << 8) <= limb; }
  // Candidates j, j + 1, ... with their test results, without a branch: each
  // goes to the next free ring slot, which only a hit keeps (a miss is
  // overwritten by the next candidate, and slots at or past the count are
  // never flushed).
  template <int K>
  __device__ void put_batch(const int (&val)[K], const bool (&hit)[K]) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      *(lds_int *)(size_t)((cb & kRingRowMask) | lane_at) = val[u];
      cb += hit[u] ? 256u : 0u;
    }
  }
  __device__ void flush_to(int upto) {  // wave-uniform
    const int n = cnt();
    for (int k = flushed; k < upto; ++k)
      if (live && k < n && k < cap && k < direct)
        list[base + (long)k * kListLanes] = ring[((unsigned)k & (kListRing - 1)) * 64 + lane];
    flushed = upto > flushed ? upto : flushed;
    if (direct == 0x7fffffff && limb != 0u)
      limb = (unsigned)(flushed + kListRing < cap ? flushed + kListRing : cap) << 8;
  }
  __device__ void row_done() {  // wave-uniform: rows every lane has completed
    int m = live ? cnt() : 0x7fffffff;
    for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
    m = m < cap ? m : cap;
    if (m > flushed) flush_to(m);
  }
  __device__ void finish() {
    const int n = cnt();
    const int m = wave_max_i32(live ? (n < cap ? n : cap) : 0);
    flush_to(m);
  }
};

// Hits of one batch of candidates (indices val, in scan order) go to the
// writer in that order (the order of a one-by-one scan): through put_batch
// while every lane's ring has room, else one by one through put.
__device__ __forceinline__ void emit_batch(ListWriter &w, const int (&val)[kListBatch],
                                           const bool (&hit)[kListBatch]) {
  if (__all(w.room(kListBatch))) {  // one wave-uniform branch in the usual case
    w.put_batch(val, hit);
  } else {
#pragma unroll
    for (int u = 0; u < kListBatch; ++u)
      if (hit[u]) w.put(val[u]);
  }
}

// Candidates j, j + 1, ... of a batch tested against p; places at or past e0
// (the padding after the array, or the next span's particles) are masked out.
// self: the index to skip (the particle itself), -1 for none.
__device__ __forceinline__ void test_batch(ListWriter &w, const float4 (&o)[kListBatch], int j,
                                           int e0, const float4 &p, float r2_lt, int self) {
  bool hit[kListBatch];
  int val[kListBatch];
#pragma unroll
  for (int u = 0; u < kListBatch; ++u) {
    val[u] = j + u;
    hit[u] = (j + u < e0) & (j + u != self) &
             (dist2(p.x, p.y, p.z, o[u].x, o[u].y, o[u].z) < r2_lt);
  }
  emit_batch(w, val, hit);
}

// Candidates [s0, e0) of one span, kListBatch at a time. The last batch may
// read up to kListBatch - 1 places past e0 (at most kPointPad past the
// array's last particle) and masks them out.
__device__ __forceinline__ void span_hits(const float4 *__restrict__ pts, int s0, int e0,
                                          const float4 &p, float r2_lt, int self, ListWriter &w) {
  for (int j = s0; j < e0; j += kListBatch) {
    float4 o[kListBatch];
#pragma unroll
    for (int u = 0; u < kListBatch; ++u) o[u] = pts[j + u];
    test_batch(w, o, j, e0, p, r2_lt, self);
  }
}

// Whether particle p can have boundary candidates: its query range lies within
// 2 cells of its own cell on every axis, and the slots within 2 cells of every
// boundary particle's cell are marked (hashing is modular in the cell
// coordinate, so a candidate always marks the query's own slot). Unknown
// marks, or coordinates where the hash clamps, count as near.
__device__ __forceinline__ bool near_boundary(const Bufs &b, const SphConst &c, const GridRef &g,
                                              const float4 &p) {
  if (!b.bnear) return true;
  const int cx = (int)(p.x * c.cell_inv), cy = (int)(p.y * c.cell_inv), cz = (int)(p.z * c.cell_inv);
  const int lim = -(1 << 20) + 8;
  if (cx < lim || cy < lim || cz < lim) return true;
  const int hx = hash_axis(cx, g.dims[0]), hy = hash_axis(cy, g.dims[1]), hz = hash_axis(cz, g.dims[2]);
  const int d0 = g.dims[0], d1 = g.dims[1];
  const int sl = c.layout == WSPH_LAYOUT_MORTON ? (int)morton3(hx, hy, hz) : hz * (d0 * d1) + hy * d0 + hx;
  return b.bnear[(g.start - b.bstart) + sl] != 0u;
}

// Largest neighbour counts of one wave of the list pass (and the overflow
// flags, rare): into slot `wave` mod kListMaxSlots.
__device__ __forceinline__ void record_list_max(const Bufs &b, int wave, int mf, int ms) {
  if ((threadIdx.x & 63) != 0) return;
  int *slot = b.list_max + (wave & (kListMaxSlots - 1)) * kListMaxStride;
  atomicMax(slot, mf);
  if (ms > 0) atomicMax(slot + 1, ms);
  if (mf > b.cap_ff) atomicOr(&b.st->status, (int)ST_FF_OVERFLOW);
  if (ms > b.cap_fs) atomicOr(&b.st->status, (int)ST_FS_OVERFLOW);
}

// 5 waves per SIMD (registers held to 96, no spills; the 128-thread blocks'
// LDS rings admit 5): 289 -> 279 us at 1M particles against the compiler's
// 100 registers and 4 waves; 6 waves spill.
__global__ void __attribute__((amdgpu_waves_per_eu(5))) k_lists(Bufs b, SphConst c) {
  int *ring = reinterpret_cast<int *>(g_smem) + (threadIdx.x >> 6) * (kListRing * 64);
  const int lane = (int)(threadIdx.x & 63);
  const int i = particle_index();
  const bool live = i < b.n;
  const float4 p = live ? b.pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const int scene = b.scenes && live ? b.sid[i] : 0;
  ListWriter wf{ring, b.nbr_ff, list_base(i, b.cap_ff), i, lane, b.cap_ff, live};
  wf.init();
  const GridRef fg = fluid_grid(b, c, scene);
  for_query_rows(
      p.x, p.y, p.z, live, c, fg.dims, fg.start, fg.count,
      [&](int s0, int e0) {
        // distance < SMOOTHING_LENGTH and index != i (IISPH.py:72)
        span_hits(b.pos, s0, e0, p, c.r2_lt, i, wf);
      },
      [&]() { wf.row_done(); });
  wf.finish();
  ListWriter wb{ring, b.nbr_fs, list_base(i, b.cap_fs), i, lane, b.cap_fs, live};
  wb.init();
  const GridRef bg = boundary_grid(b, c, scene);
  if (b.nb > 0 && __ballot(live && near_boundary(b, c, bg, p))) {
    for_query_rows(
        p.x, p.y, p.z, live, c, bg.dims, bg.start, bg.count,
        [&](int s0, int e0) {
          // distance < SMOOTHING_LENGTH (IISPH.py:129)
          span_hits(b.bpos, s0, e0, p, c.r2_lt, -1, wb);
        },
        [&]() { wb.row_done(); });
    wb.finish();
  }
  const int nf = wf.cnt(), ns = wb.cnt();
  if (live) {
    b.cnt_ff[i] = nf < b.cap_ff ? nf : b.cap_ff;
    b.cnt_fs[i] = ns < b.cap_fs ? ns : b.cap_fs;
  }
  record_list_max(b, i / kListLanes, wave_max_i32(nf), wave_max_i32(ns));
}

// Marks of near_boundary: every slot within 2 cells (each axis) of a boundary
// particle's cell.
__global__ void k_mark_bnear(Bufs b, SphConst c, unsigned *__restrict__ bnear) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.nb) return;
  const float4 p = b.bpos[k];
  const GridRef g = boundary_grid(b, c, b.scenes ? b.bsid[k] : 0);
  const long base = g.start - b.bstart;
  const int d0 = g.dims[0], d1 = g.dims[1], d2 = g.dims[2];
  const int cx = (int)(p.x * c.cell_inv), cy = (int)(p.y * c.cell_inv), cz = (int)(p.z * c.cell_inv);
  for (int dz = -2; dz <= 2; ++dz) {
    const int hz = hash_axis(cz + dz, d2);
    for (int dy = -2; dy <= 2; ++dy) {
      const int hy = hash_axis(cy + dy, d1);
      for (int dx = -2; dx <= 2; ++dx) {
        const int hx = hash_axis(cx + dx, d0);
        const int sl = c.layout == WSPH_LAYOUT_MORTON ? (int)morton3(hx, hy, hz)
                                                      : hz * (d0 * d1) + hy * d0 + hx;
        bnear[base + sl] = 1u;
      }
    }
  }
}

// ---- compute_boundary_density, IISPH.py:23-47 --------------------------
template <bool kLds>
__global__ void k_boundary_psi(Bufs b, SphConst c) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  if (i >= b.nb) return;
  const float4 p = b.bpos[i];
  float rho = W_OF(T, 0.0f);
  const GridRef bg = boundary_grid(b, c, b.scenes ? b.bsid[i] : 0);
  for_query_spans(p.x, p.y, p.z, c, bg.dims, bg.start, bg.count, [&](int s0, int e0) {
    for_span(b.bpos, s0, e0, [&](int u, const float4 &o) {
      const float d2 = dist2(p.x, p.y, p.z, o.x, o.y, o.z);
      if (d2 < c.r2_lt && u != i) rho += W_OF(T, sqrtf(d2));
    });
  });
  reinterpret_cast<float *>(&b.bpos[i])[3] = c.rho0 / rho;
}

// v_adv of predict_v_adv (IISPH.py:217-278) with the Laplacian term switched
// off (VIS_MU = 0): v + ((0 * vis_mu) / rho + g) dt.
__device__ __forceinline__ float4 vadv_plain(float4 v, float rho, float dt, const SphConst &c) {
  const float gx = 0.0f, gy = c.g_y, gz = 0.0f;
  const float lx = 0.0f, ly = 0.0f, lz = 0.0f;
  return make_float4(v.x + (lx * c.vis_mu / rho + gx) * dt, v.y + (ly * c.vis_mu / rho + gy) * dt,
                     v.z + (lz * c.vis_mu / rho + gz) * dt, 0.0f);
}

// ---- compute_density, IISPH.py:177-214 -----------------------------------
// Also the gradient sums of d_ii (compute_term_d, IISPH.py:359-382), which need
// only positions, psi and rho_i: m sum_j -grad W_ij + sum_b -psi_b grad W_ib
// times dt^2 / rho_i^2, accumulated in list order in the same loops and left
// in sap for k_assemble. One rounded square root serves W and grad W: the
// radicand of wp.length(x_i - x_j) and of |x_j - x_i| is the same float.
template <bool kLds>
__global__ void k_density(Bufs b, SphConst c, int fuse_vadv) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  if (i >= b.n) return;
  const float4 p = b.pos[i];
  const float4 *__restrict__ pos = b.pos;
  const float4 *__restrict__ bpos = b.bpos;
  float t1 = 0.0f, t2 = 0.0f;
  float dx1 = 0.0f, dy1 = 0.0f, dz1 = 0.0f, dx2 = 0.0f, dy2 = 0.0f, dz2 = 0.0f;
  for_neighbors<kBatch>(
      b.nbr_ff, b.cap_ff, i, b.cnt_ff[i], [&](int j, int) { return ld_at(pos, j); },
      [&](const float4 &o) {
        const float3 r = sub3(o, p);
        const float len = sqrt_rn(r.x * r.x + r.y * r.y + r.z * r.z);
        t1 += W_OF(T, len);
        const float3 gw = grad_w_len(T.g, c.ng, c.g_scale, len, r);
        dx1 = dx1 - gw.x;
        dy1 = dy1 - gw.y;
        dz1 = dz1 - gw.z;
      });
  for_neighbors<kBatch>(
      b.nbr_fs, b.cap_fs, i, b.cnt_fs[i], [&](int j, int) { return ld_at(bpos, j); },
      [&](const float4 &o) {
        const float3 r = sub3(o, p);
        const float len = sqrt_rn(r.x * r.x + r.y * r.y + r.z * r.z);
        t2 += W_OF(T, len) * o.w;
        const float3 gw = grad_w_len(T.g, c.ng, c.g_scale, len, r);
        dx2 = dx2 - gw.x * o.w;
        dy2 = dy2 - gw.y * o.w;
        dz2 = dz2 - gw.z * o.w;
      });
  const float rho = t1 * c.mass + t2 + W_OF(T, 0.0f) * c.mass;
  reinterpret_cast<float *>(&b.pos[i])[3] = rho;
  const float dt = (float)b.st->dt;
  const float dt2 = dt * dt;
  const float ri2 = rho * rho;
  b.sap[i] = make_float4((dx1 * c.mass + dx2) * dt2 / ri2, (dy1 * c.mass + dy2) * dt2 / ri2,
                         (dz1 * c.mass + dz2) * dt2 / ri2, 0.0f);
  if (fuse_vadv) b.vadv[i] = vadv_plain(b.vel[i], rho, dt, c);
}

// ---- predict_v_adv, IISPH.py:217-278 (+ XSPH smoothing) ------------------
template <bool kLds>
__global__ void k_vadv(Bufs b, SphConst c) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  if (i >= b.n) return;
  const float dt = (float)b.st->dt;
  const float4 p = b.pos[i];
  const float4 vi = b.vel[i];
  const int nf = b.cnt_ff[i], ns = b.cnt_fs[i];
  const float4 *__restrict__ pos = b.pos;
  const float4 *__restrict__ vel = b.vel;
  const float4 *__restrict__ bpos = b.bpos;
  float lx = 0.0f, ly = 0.0f, lz = 0.0f;
  if (c.vis_mu != 0.0f) {
    float ax = 0.0f, ay = 0.0f, az = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
    for_neighbors<kBatch>(
        b.nbr_ff, b.cap_ff, i, nf, [&](int j, int) { return F4x2{ld_at(pos, j), ld_at(vel, j)}; },
        [&](const F4x2 &v) {
          const float4 o = v.a, vj = v.b;
          const float3 xij = sub3(o, p);
          const float3 gw = GW_OF(T, xij);
          const float3 vd = make_float3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z);
          const float dd = dot3(vd, xij);
          const float d = dist3(p.x, p.y, p.z, o.x, o.y, o.z);
          const float r2 = d * d;
          ax = ax - gw.x * dd / o.w / r2;
          ay = ay - gw.y * dd / o.w / r2;
          az = az - gw.z * dd / o.w / r2;
        });
    for_neighbors<kBatch>(
        b.nbr_fs, b.cap_fs, i, ns, [&](int j, int) { return ld_at(bpos, j); },
        [&](const float4 &va_) {
          const F4x2 v{va_, va_};
          const float4 o = v.a;
          const float3 xib = sub3(o, p);
          const float3 gw = GW_OF(T, xib);
          const float dd = dot3(make_float3(vi.x, vi.y, vi.z), xib);
          const float d = dist3(p.x, p.y, p.z, o.x, o.y, o.z);
          const float r2 = d * d;
          bx = bx - gw.x * o.w * dd / r2;
          by = by - gw.y * o.w * dd / r2;
          bz = bz - gw.z * o.w * dd / r2;
        });
    lx = 10.0f * (ax * c.mass + bx / c.rho0);
    ly = 10.0f * (ay * c.mass + by / c.rho0);
    lz = 10.0f * (az * c.mass + bz / c.rho0);
  }
  float vx = vi.x + (lx * c.vis_mu / p.w + 0.0f) * dt;
  float vy = vi.y + (ly * c.vis_mu / p.w + c.g_y) * dt;
  float vz = vi.z + (lz * c.vis_mu / p.w + 0.0f) * dt;
  if (c.xsph != 0.0f) {
    // XSPH: v_adv += xsph * sum_j (v_j - v_i) * ((m / rho_j) * W_ij)
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for_neighbors<kBatch>(
        b.nbr_ff, b.cap_ff, i, nf, [&](int j, int) { return F4x2{ld_at(pos, j), ld_at(vel, j)}; },
        [&](const F4x2 &v) {
          const float4 o = v.a, vj = v.b;
          const float coef =
              div_by(c.mass, o.w, rcp_rn(o.w)) * W_OF(T, dist3(p.x, p.y, p.z, o.x, o.y, o.z));
          ax = ax + (vj.x - vi.x) * coef;
          ay = ay + (vj.y - vi.y) * coef;
          az = az + (vj.z - vi.z) * coef;
        });
    vx = vx + ax * c.xsph;
    vy = vy + ay * c.xsph;
    vz = vz + az * c.xsph;
  }
  b.vadv[i] = make_float4(vx, vy, vz, 0.0f);
}

// ---- predict_rho_adv (IISPH.py:281-329), compute_term_d (IISPH.py:332-382)
//      and compute_term_a (IISPH.py:385-428) in one pass ---------------------
template <bool kLds>
__global__ void k_assemble(Bufs b, SphConst c, int start_mode) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  if (i >= b.n) return;
  const float dt = (float)b.st->dt;
  const float dt2 = dt * dt;
  const float t3 = c.mass * dt2;
  const float4 p = b.pos[i];
  const float4 va = b.vadv[i];
  const float ri2 = p.w * p.w;
  const float yi = 1.0f / ri2;
  const int nf = b.cnt_ff[i], ns = b.cnt_fs[i];
  const float4 *__restrict__ pos = b.pos;
  const float4 *__restrict__ vadv = b.vadv;
  const float4 *__restrict__ bpos = b.bpos;
  const float4 *__restrict__ bvel = b.bvel;
  // d_ii from k_density; rho_adv (predict_rho_adv) and a_ii (compute_term_a)
  // share the neighbour loops, each sum in list order
  const float4 di = b.sap[i];
  const float dix = di.x, diy = di.y, diz = di.z;
  float tr1 = 0.0f, tr2 = 0.0f, ta1 = 0.0f, ta2 = 0.0f;
  for_neighbors<kBatch>(
      b.nbr_ff, b.cap_ff, i, nf, [&](int j, int) { return F4x2{ld_at(pos, j), ld_at(vadv, j)}; },
      [&](const F4x2 &v) {
        const float3 gw = GW_OF(T, sub3(v.a, p));
        tr1 += dot3(make_float3(va.x - v.b.x, va.y - v.b.y, va.z - v.b.z), gw);
        const float3 dd = make_float3(dix - div_by(gw.x * t3, ri2, yi),
                                      diy - div_by(gw.y * t3, ri2, yi),
                                      diz - div_by(gw.z * t3, ri2, yi));
        ta1 += dot3(dd, gw);
      });
  for_neighbors<kBatch>(
      b.nbr_fs, b.cap_fs, i, ns, [&](int j, int) { return F4x2{ld_at(bpos, j), ld_at(bvel, j)}; },
      [&](const F4x2 &v) {
        const float4 o = v.a, vb = v.b;
        const float3 gw = GW_OF(T, sub3(o, p));
        tr2 += dot3(make_float3(va.x - vb.x, va.y - vb.y, va.z - vb.z), gw) * o.w;
        ta2 += dot3(make_float3(dix, diy, diz), gw) * o.w;
      });
  const float rho_adv = p.w + (tr1 * c.mass + tr2) * dt;
  const float a_ii = ta1 * c.mass + ta2;
  b.aux[i] = make_float4(rho_adv, a_ii, 0.0f, 0.0f);
  float pv;
  if (start_mode == WSPH_WARM_ZERO)
    pv = update_p(0.0f, rho_adv, 0.0f, a_ii, c);  // A 0 = 0: the first Jacobi update
  else if (start_mode == WSPH_WARM_PREVIOUS || start_mode == WSPH_WARM_LEARNED)
    pv = b.pprev[i];  // the learned start overwrites it in the feature pass
  else
    pv = b.pstart[i];
  if (b.seed)  // self features of the pressure predictor
    b.seed[i] = make_float4(clamp_feat((rho_adv / c.rho0 - 1.0f) * 100.0f),
                            clamp_feat(a_ii * c.feat_k1 / (dt * dt)),
                            clamp_feat((float)nf * (1.0f / 32.0f)),
                            clamp_feat(b.pprev[i] * c.inv_pref));
  b.dp[0][i] = make_float4(dix, diy, diz, pv);
  b.sap[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  b.ap[i] = 0.0f;
}

// ---- compute_term_Ap_1, IISPH.py:431-466 ---------------------------------

template <typename TT>
__device__ __forceinline__ void j1_particle(const Bufs &b, const SphConst &c, const TT &T, int i,
                                            int src, float t3) {
  const float4 p = b.pos[i];
  const float4 *__restrict__ dp = b.dp[src];
  const float4 *__restrict__ pos = b.pos;
  const float pi = dp[i].w;
  const float ri2 = p.w * p.w;
  const float yi = 1.0f / ri2;
  float t1 = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for_neighbors<kBatchJ1>(
      b.nbr_ff, b.cap_ff, i, b.cnt_ff[i], [&](int j, int) { return F4x2{ld_at(pos, j), ld_at(dp, j)}; },
      [&](const F4x2 &v) {
        const float4 o = v.a, dj = v.b;
        const float3 gw = GW_OF(T, sub3(o, p));
        // d_ji = grad W_ij * t3 / rho_i^2, d_ij = -grad W_ij * t3 / rho_j^2 (IISPH.py:372-373);
        // both divide a = grad W_ij * t3, and (-a) / b = -(a / b) exactly
        const float ax = gw.x * t3, ay = gw.y * t3, az = gw.z * t3;
        const float3 term = make_float3(div_by(ax, ri2, yi) * pi - dj.x * dj.w,
                                        div_by(ay, ri2, yi) * pi - dj.y * dj.w,
                                        div_by(az, ri2, yi) * pi - dj.z * dj.w);
        t1 += dot3(term, gw);
        const float rj2 = o.w * o.w;
        const float yj = rcp_rn(rj2);
        sx = sx - div_by(ax, rj2, yj) * dj.w;
        sy = sy - div_by(ay, rj2, yj) * dj.w;
        sz = sz - div_by(az, rj2, yj) * dj.w;
      });
  b.sap[i] = make_float4(sx, sy, sz, t1 * c.mass);
}

// ---- compute_term_Ap_2 (IISPH.py:469-518), update_rho_error (IISPH.py:542-568)
//      and the next update_p (IISPH.py:521-539) ----------------------------
template <typename TT>
__device__ __forceinline__ void j2_particle(const Bufs &b, const SphConst &c, const TT &T, int i,
                                            int src, unsigned long long &rsum, unsigned &rcnt) {
  const float4 p = b.pos[i];
  const float4 si = b.sap[i];
  const float4 *__restrict__ pos = b.pos;
  const float4 *__restrict__ sap = b.sap;
  const float4 *__restrict__ bpos = b.bpos;
  float t1 = 0.0f, t2 = 0.0f;
  for_neighbors<kBatchJ2>(
      b.nbr_ff, b.cap_ff, i, b.cnt_ff[i], [&](int j, int) { return F4x2{ld_at(pos, j), ld_at(sap, j)}; },
      [&](const F4x2 &v) {
        const float3 gw = GW_OF(T, sub3(v.a, p));
        t1 += dot3(make_float3(si.x - v.b.x, si.y - v.b.y, si.z - v.b.z), gw);
      });
  for_neighbors<kBatchJ2>(
      b.nbr_fs, b.cap_fs, i, b.cnt_fs[i], [&](int j, int) { return ld_at(bpos, j); },
      [&](const float4 &va_) {
        const F4x2 v{va_, va_};
        const float3 gw = GW_OF(T, sub3(v.a, p));
        t2 += dot3(make_float3(si.x, si.y, si.z), gw) * v.a.w;
      });
  const float4 di = b.dp[src][i];
  const float4 au = b.aux[i];
  const float ap = si.w + (t1 * c.mass + t2 + au.y * di.w);
  b.ap[i] = ap;
  const float ur = au.x + ap;
  if (ur > c.rho0) {
    rsum += residual_fixed(fabsf(ur / c.rho0 - 1.0f));
    rcnt += 1u;
  }
  b.dp[src ^ 1][i] = make_float4(di.x, di.y, di.z, update_p(di.w, au.x, ap, au.y, c));
}

__device__ __forceinline__ void block_residual_add(DevState *st, int slot,
                                                   unsigned long long rsum, unsigned rcnt) {
  __shared__ unsigned long long s_sum[16];
  __shared__ unsigned s_cnt[16];
  rsum = wave_sum_u64(rsum);
  rcnt = wave_sum_u32(rcnt);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_sum[wid] = rsum;
    s_cnt[wid] = rcnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long S = 0;
    unsigned Cn = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      S += s_sum[w];
      Cn += s_cnt[w];
    }
    if (Cn) {
      atomicAdd(&st->res_sum[slot], S);
      atomicAdd(&st->res_cnt[slot], Cn);
    }
  }
  __syncthreads();
}

// Block partial of the residual sum and count (no atomics: one slot per block).
__device__ __forceinline__ void block_residual_store(const Bufs &b, unsigned long long rsum,
                                                     unsigned rcnt) {
  __shared__ unsigned long long s_sum[16];
  __shared__ unsigned s_cnt[16];
  rsum = wave_sum_u64(rsum);
  rcnt = wave_sum_u32(rcnt);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_sum[wid] = rsum;
    s_cnt[wid] = rcnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long S = 0;
    unsigned Cn = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      S += s_sum[w];
      Cn += s_cnt[w];
    }
    b.res_part[blockIdx.x] = S;
    b.cnt_part[blockIdx.x] = Cn;
  }
}

// Sum of the block partials of iteration `slot` into DevState (the host-driven
// loop, which then reads the sums and runs the stop test itself).
__global__ void k_residual_reduce(Bufs b, int nparts, int slot) {
  __shared__ unsigned long long s_sum[16];
  __shared__ unsigned s_cnt[16];
  DevState *st = b.st;
  unsigned long long S = 0;
  unsigned Cn = 0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) {
    S += b.res_part[q];
    Cn += b.cnt_part[q];
  }
  S = wave_sum_u64(S);
  Cn = wave_sum_u32(Cn);
  if ((threadIdx.x & 63) == 0) {
    s_sum[threadIdx.x >> 6] = S;
    s_cnt[threadIdx.x >> 6] = Cn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    S = 0;
    Cn = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      S += s_sum[w];
      Cn += s_cnt[w];
    }
    st->res_sum[slot] = S;
    st->res_cnt[slot] = Cn;
  }
}

// ---- batched scenes: per-scene residual and stop test ----------------------
// Particle i takes part in the Jacobi iteration unless its scene has stopped.
__device__ __forceinline__ bool scene_active(const Bufs &b, int i, int &scene) {
  scene = -1;
  if (i >= b.n) return false;
  if (!b.scenes) return true;
  scene = b.sid[i];
  return !b.sc_done[scene];
}

// Residual terms of the wave's particles added to their scenes' sums of
// iteration k (integer atomics: the sums do not depend on order). Lanes are
// grouped by scene; particles are sorted scene-major, so a wave usually
// holds one scene and issues one atomic pair. Every lane of the wave calls
// this; scene < 0 contributes nothing.
__device__ __forceinline__ void scene_residual_add(const Bufs &b, int k, int scene,
                                                   unsigned long long rs, unsigned rc) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(scene >= 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int sl = __shfl(scene, leader, 64);
    const bool mine = scene == sl;
    const unsigned long long v = wave_sum_u64(mine ? rs : 0ull);
    const unsigned cnum = wave_sum_u32(mine ? rc : 0u);
    if (lane == leader && cnum) {
      atomicAdd(&b.sc_sum[(long)k * b.nscenes + sl], v);
      atomicAdd(&b.sc_cnt[(long)k * b.nscenes + sl], cnum);
    }
    pending &= ~__ballot(mine);
  }
}

// Stop test of every running scene after iteration k (IISPH.py:1200-1207 per
// scene); `done` receives the decisions. Returns 1 if any scene continues.
// Results are written when `write` is set.
__device__ __forceinline__ int batch_decide(const Bufs &b, int k, int src_k, int min_it,
                                            int max_it, int force_it, double eta,
                                            unsigned char *done, bool write) {
  int running = 0;
  for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x) {
    if (done[sc]) continue;
    const unsigned long long S = __hip_atomic_load(&b.sc_sum[(long)k * b.nscenes + sc],
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned Cn = __hip_atomic_load(&b.sc_cnt[(long)k * b.nscenes + sc],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double e;
    if (jacobi_continue(k, S, Cn, min_it, max_it, force_it, eta, &e)) {
      running = 1;
    } else {
      done[sc] = 1;
      if (write) {
        b.sc_iters[sc] = k + 1;
        b.sc_res[sc] = e;
        b.sc_final[sc] = src_k;
      }
    }
  }
  return __syncthreads_or(running);
}

// One block: the per-scene stop test of a batched device-driven iteration.
__global__ void k_batch_stop(Bufs b, int src0, int min_it, int max_it, int force_it, double eta) {
  __shared__ unsigned char s_done[kMaxScenes];
  __shared__ double s_emax[16];
  __shared__ int s_conv[16];
  DevState *st = b.st;
  if (st->done) return;
  const int k = st->iter;
  for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x) s_done[sc] = (unsigned char)b.sc_done[sc];
  __syncthreads();
  const int any = batch_decide(b, k, src0 ^ (k & 1), min_it, max_it, force_it, eta, s_done, true);
  for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x) b.sc_done[sc] = s_done[sc];
  if (any) {
    if (threadIdx.x == 0) st->iter = k + 1;
    return;
  }
  __syncthreads();
  // every scene has stopped: step results (largest residual, all converged)
  double emax = 0.0;
  int conv = 1;
  for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x) {
    const double e = b.sc_res[sc];
    emax = e > emax ? e : emax;
    conv &= e <= eta;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(emax, o, 64);
    emax = t > emax ? t : emax;
    conv &= __shfl_xor(conv, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_emax[threadIdx.x >> 6] = emax;
    s_conv[threadIdx.x >> 6] = conv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      emax = s_emax[w] > emax ? s_emax[w] : emax;
      conv &= s_conv[w];
    }
    st->done = 1;
    st->iterations = k + 1;
    st->ap_evals += k + 1;
    st->residual = emax;
    st->converged = conv;
    st->final_src = src0 ^ (k & 1);
    st->done_tag = st->step_tag;
    if (b.done_host) *(volatile int *)b.done_host = st->step_tag;
  }
}

// Scenes that stopped before the last iteration hold their final iterate in
// dp[sc_final]; copy it to dp[final_src].
__global__ void k_batch_finalize(Bufs b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const int f = b.st->final_src;
  const int mine = b.sc_final[b.sid[i]];
  if (mine != f) b.dp[f][i] = b.dp[mine][i];
}

// Per-step reset of the scenes' stop state and of the residual slots of the
// iterations the step may run.
__global__ void k_batch_reset(Bufs b, int slots) {
  const long total = (long)slots * b.nscenes;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long)gridDim.x * blockDim.x) {
    b.sc_sum[q] = 0ull;
    b.sc_cnt[q] = 0u;
    if (q < b.nscenes) {
      b.sc_done[q] = 0;
      b.sc_iters[q] = 0;
      b.sc_final[q] = 0;
      b.sc_res[q] = 0.0;
    }
  }
}

template <bool kLds>
__global__ void k_j1(Bufs b, SphConst c, int src) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  const float dt = (float)b.st->dt;
  const float t3 = c.mass * (dt * dt);
  if (i < b.n) j1_particle(b, c, T, i, src, t3);
}

template <bool kLds>
__global__ void k_j2(Bufs b, SphConst c, int src, int slot) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const int i = particle_index();
  unsigned long long rs = 0;
  unsigned rc = 0;
  if (i < b.n) j2_particle(b, c, T, i, src, rs, rc);
  if (slot >= 0) block_residual_store(b, rs, rc);
  if (slot == -1 && blockIdx.x == 0 && threadIdx.x == 0) b.st->ap_evals += 1;  // the A p0 evaluation
}

// ---- device-resident Jacobi loop control ----------------------------------
// Progress of the loop for the host, which paces its enqueues on it: the
// step's tag and the iterations whose stop test has run, in one 64-bit store
// to pinned host memory. J1 of iteration k > 0 stores it (k tests have run),
// so the store's round trip over the host link overlaps J1's own work
// instead of delaying the end of the one-block stop kernel.
__device__ __forceinline__ void report_progress(const Bufs &b, const DevState *st, int iters) {
  if (b.prog_host)
    *(volatile unsigned long long *)b.prog_host =
        ((unsigned long long)(unsigned)st->step_tag << 32) | (unsigned)iters;
}

// Loop end of the device-driven Jacobi loop after iteration k (IISPH.py:1200-1207):
// the iterate of iteration k (buffer src0 ^ (k & 1)) is final; its update,
// already written to the other buffer, is discarded.
__device__ __forceinline__ void record_loop_end(const Bufs &b, DevState *st, int k, double e,
                                                int src0, double eta) {
  st->done = 1;
  st->iterations = k + 1;
  st->ap_evals += k + 1;
  st->residual = e;
  st->converged = e <= eta;
  st->final_src = src0 ^ (k & 1);
  st->done_tag = st->step_tag;
  if (b.done_host) *(volatile int *)b.done_host = st->step_tag;
}

// Residual sum and count of the J2 block partials (res_part / cnt_part,
// nparts of them), summed by every thread of the calling block (integer sums:
// exact, any order).
__device__ __forceinline__ void block_sum_partials(const Bufs &b, int nparts,
                                                   unsigned long long &S, unsigned &Cn) {
  __shared__ unsigned long long s_sum[16];
  __shared__ unsigned s_cnt[16];
  S = 0;
  Cn = 0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) {
    S += b.res_part[q];
    Cn += b.cnt_part[q];
  }
  S = wave_sum_u64(S);
  Cn = wave_sum_u32(Cn);
  if ((threadIdx.x & 63) == 0) {
    s_sum[threadIdx.x >> 6] = S;
    s_cnt[threadIdx.x >> 6] = Cn;
  }
  __syncthreads();
  S = 0;
  Cn = 0;
  for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
    S += s_sum[w];
    Cn += s_cnt[w];
  }
}

// Each iteration k is J1 and J2. One scene: J2's blocks store their residual
// partials, and J1 of iteration k + 1 runs the stop test of iteration k: every
// block sums the same partials and takes the same decision, and block 0
// records the loop's end. The loop's last iteration is tested by
// k_jacobi_last_test. Batched scenes keep a one-block stop kernel per
// iteration (k_batch_stop), which advances DevState::iter. Kernels return at
// once once the loop has finished, so the host can enqueue iterations ahead
// without knowing when the loop stops.
template <bool kLds>
__global__ void k_j1_dev(Bufs b, SphConst c, int src0, int k, LoopCtl L) {
  DevState *st = b.st;
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (b.scenes) {
    k = st->iter;
  } else if (k > 0) {
    unsigned long long S;
    unsigned Cn;
    block_sum_partials(b, (int)gridDim.x, S, Cn);
    double e;
    if (!jacobi_continue(k - 1, S, Cn, L.min_it, L.max_it, L.force_it, L.eta, &e)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->res_sum[k - 1] = S;
        st->res_cnt[k - 1] = Cn;
        record_loop_end(b, st, k - 1, e, src0, L.eta);
      }
      return;
    }
  }
  const Tabs T = stage_tabs<kLds>(b, c);
  if (k > 0 && blockIdx.x == 0 && threadIdx.x == 0) report_progress(b, st, k);
  const int i = particle_index();
  const float dt = (float)st->dt;
  const float t3 = c.mass * (dt * dt);
  int scene;
  if (scene_active(b, i, scene)) j1_particle(b, c, T, i, src0 ^ (k & 1), t3);
}

// J2 of a device-driven iteration k (batched scenes: DevState::iter).
template <bool kLds>
__global__ void k_j2_dev(Bufs b, SphConst c, int src0, int k) {
  if (__hip_atomic_load(&b.st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const Tabs T = stage_tabs<kLds>(b, c);
  if (b.scenes) k = b.st->iter;
  const int i = particle_index();
  unsigned long long rs = 0;
  unsigned rc = 0;
  int scene;
  const bool act = scene_active(b, i, scene);
  if (act) j2_particle(b, c, T, i, src0 ^ (k & 1), rs, rc);
  if (b.scenes)
    scene_residual_add(b, k, act ? scene : -1, rs, rc);
  else
    block_residual_store(b, rs, rc);
}

// Stop test of the last queued iteration k of a one-scene device loop (the
// test of every earlier iteration runs in the next J1).
__global__ void k_jacobi_last_test(Bufs b, int nparts, int src0, int k, LoopCtl L) {
  DevState *st = b.st;
  if (st->done) return;
  unsigned long long S;
  unsigned Cn;
  block_sum_partials(b, nparts, S, Cn);
  double e;
  if (threadIdx.x == 0 &&
      !jacobi_continue(k, S, Cn, L.min_it, L.max_it, L.force_it, L.eta, &e)) {
    st->res_sum[k] = S;
    st->res_cnt[k] = Cn;
    record_loop_end(b, st, k, e, src0, L.eta);
  }
}

// ---- kick (IISPH.py:674-688) and drift (IISPH.py:691-699) ----------------
// Also keeps the final pressure (pfinal) and the pressure carried to the
// next step's start value (pprev): zero where a_ii fell inside the guard band
// of update_p (|a_ii| < INV_SMALL, IISPH.py:535), because the pressure of such
// a weakly coupled particle grows without bound and means nothing.
// float -> unsigned with the same order (for max via integer max)
__device__ __forceinline__ unsigned ordered_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_integrate(Bufs b, SphConst c, int src) {
  __shared__ unsigned s_vmax[16];
  __shared__ int s_bad[16];
  __shared__ double s_ke[16], s_y[16], s_en[16], s_ec[16];
  __shared__ unsigned s_ymax[16];
  double ke = 0.0, ysum = 0.0, en = 0.0, ec = 0.0;
  unsigned ymax = 0u;
  if (src < 0) src = b.st->final_src;
  const int i = particle_index();
  unsigned vbits = 0;
  int bad = 0;
  if (i < b.n) {
    const double dtd = b.st->dt;
    const float dt = (float)dtd;
    const float inv_dt = (float)(1.0 / dtd);
    const float4 di = b.dp[src][i];
    const float4 s = b.sap[i];
    const float4 va = b.vadv[i];
    const float vx = va.x + inv_dt * (di.x * di.w + s.x);
    const float vy = va.y + inv_dt * (di.y * di.w + s.y);
    const float vz = va.z + inv_dt * (di.z * di.w + s.z);
    const float l = sqrtf(vx * vx + vy * vy + vz * vz);
    vbits = __float_as_uint(l);
    const float4 p = b.pos[i];
    const float x = p.x + vx * dt, y = p.y + vy * dt, z = p.z + vz * dt;
    b.pos[i] = make_float4(x, y, z, p.w);
    b.vel[i] = make_float4(vx, vy, vz, 0.0f);
    b.pfinal[i] = di.w;
    const float4 au = b.aux[i];
    const float a = au.y;
    if (b.target || b.arbiter) {
      const float carry = b.pprev[i];  // this step's carried start value
      if (b.target) b.target[i] = clamp_feat((di.w - carry) * c.inv_pref);
      if (b.arbiter && b.st->net_eval) {
        // start errors weighted by min(|a_ii / a_ref|, 1): a particle's share
        // of the density residual scales with a_ii
        float wgt = fabsf(a * c.feat_k1 / (dt * dt));
        wgt = wgt < 1.0f ? wgt : 1.0f;
        const double e1 = (double)wgt * ((double)au.z - (double)di.w);
        const double e2 = (double)wgt * ((double)carry - (double)di.w);
        en = e1 * e1;
        ec = e2 * e2;
      }
    }
    b.pprev[i] = (a > -c.inv_small && a < c.inv_small) ? 0.0f : di.w;
    bad = !(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(l));
    if (bad) vbits = 0;
    ke = 0.5 * (double)c.mass * ((double)vx * vx + (double)vy * vy + (double)vz * vz);
    ysum = (double)y;
    ymax = ordered_bits(y);
  }
  vbits = wave_max_u32(vbits);
  bad = wave_max_i32(bad);
  ymax = wave_max_u32(ymax);
  for (int o = 32; o > 0; o >>= 1) {
    ke += __shfl_xor(ke, o, 64);
    ysum += __shfl_xor(ysum, o, 64);
    if (b.arbiter) {
      en += __shfl_xor(en, o, 64);
      ec += __shfl_xor(ec, o, 64);
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_vmax[wid] = vbits;
    s_bad[wid] = bad;
    s_ke[wid] = ke;
    s_y[wid] = ysum;
    s_ymax[wid] = ymax;
    s_en[wid] = en;
    s_ec[wid] = ec;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = 0, my = 0;
    int bb = 0;
    double k2 = 0.0, y2 = 0.0, n2 = 0.0, c2 = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      m = max(m, s_vmax[w]);
      my = max(my, s_ymax[w]);
      bb |= s_bad[w];
      k2 += s_ke[w];
      y2 += s_y[w];
      n2 += s_en[w];
      c2 += s_ec[w];
    }
    b.vmax_part[blockIdx.x] = m;
    b.ymax_part[blockIdx.x] = my;
    b.ke_part[blockIdx.x] = k2;
    b.y_part[blockIdx.x] = y2;
    if (b.arbiter) {
      b.err_part[2 * blockIdx.x] = n2;
      b.err_part[2 * blockIdx.x + 1] = c2;
    }
    if (bb) atomicOr(&b.st->status, (int)ST_NAN);
  }
}

// max |v| and the step statistics from the per-block partials of
// k_integrate (summed in block order: deterministic), appended to the
// history ring.
__global__ void k_vmax_reduce(Bufs b, int nparts, int start_mode, StepEnd end) {
  __shared__ unsigned s[16], sy[16];
  __shared__ int sf[16], sb[16];
  __shared__ double sk[16], ss[16], sn[16], sc[16];
  unsigned m = 0, my = 0;
  double ke = 0.0, ys = 0.0, en = 0.0, ec = 0.0;
  int lf = 0, lb = 0;  // largest neighbour counts of the step's list pass
  for (int k = threadIdx.x; k < kListMaxSlots; k += blockDim.x) {
    lf = max(lf, b.list_max[k * kListMaxStride]);
    lb = max(lb, b.list_max[k * kListMaxStride + 1]);
  }
  lf = wave_max_i32(lf);
  lb = wave_max_i32(lb);
  // fixed assignment of partials to threads and a fixed combining tree:
  // the same sum on every run
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    m = max(m, b.vmax_part[k]);
    my = max(my, b.ymax_part[k]);
    ke += b.ke_part[k];
    ys += b.y_part[k];
    if (b.arbiter) {
      en += b.err_part[2 * k];
      ec += b.err_part[2 * k + 1];
    }
  }
  m = wave_max_u32(m);
  my = wave_max_u32(my);
  for (int o = 32; o > 0; o >>= 1) {
    ke += __shfl_xor(ke, o, 64);
    ys += __shfl_xor(ys, o, 64);
    en += __shfl_xor(en, o, 64);
    ec += __shfl_xor(ec, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s[threadIdx.x >> 6] = m;
    sy[threadIdx.x >> 6] = my;
    sf[threadIdx.x >> 6] = lf;
    sb[threadIdx.x >> 6] = lb;
    sk[threadIdx.x >> 6] = ke;
    ss[threadIdx.x >> 6] = ys;
    sn[threadIdx.x >> 6] = en;
    sc[threadIdx.x >> 6] = ec;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    m = 0;
    my = 0;
    ke = 0.0;
    ys = 0.0;
    en = 0.0;
    ec = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
      m = max(m, s[w]);
      my = max(my, sy[w]);
      lf = max(lf, sf[w]);
      lb = max(lb, sb[w]);
      ke += sk[w];
      ys += ss[w];
      en += sn[w];
      ec += sc[w];
    }
    DevState *st = b.st;
    st->vmax_bits = m;
    st->max_ff = max(st->max_ff, lf);
    st->max_fs = max(st->max_fs, lb);
    StepRecord &r = b.history[st->step_count % kHistory];
    r.sim_time = st->sim_time + st->dt;
    r.residual = st->residual;
    r.kinetic = ke;
    r.mean_y = b.n > 0 ? ys / b.n : 0.0;
    r.max_y = b.n > 0 ? from_ordered_bits(my) : 0.0f;
    r.v_max = __uint_as_float(m);
    r.iterations = st->iterations;
    r.converged = st->converged;
    r.start_source = start_mode;
    r.err_net = 0.0;
    r.err_carry = 0.0;
    if (start_mode == WSPH_WARM_LEARNED) {
      // net_off still holds the choice this step's feature pass made
      if (st->net_off) r.start_source = WSPH_WARM_PREVIOUS;
      if (b.arbiter && b.n > 0 && st->net_eval) {
        r.err_net = sqrt(en / b.n);
        r.err_carry = sqrt(ec / b.n);
        // the carried pressure is compared with the network for the next step
        st->net_off = en > ec ? 1 : 0;
      }
    }
    st->step_count += 1;
    // sim_time += dt; CFL dt update (IISPH.py:1367-1371)
    const double dt = st->dt;
    st->sim_time = st->sim_time + dt;
    if (end.dt_mode == WSPH_DT_CFL) {
      const double vmax = (double)__uint_as_float(st->vmax_bits);
      const double mv = vmax > end.inv_small ? vmax : end.inv_small;
      const double nd = end.cfl_d / mv;
      st->dt = nd < end.dt_max ? nd : end.dt_max;
    }
    // cumulative counters
    long long *tot = end.totals;
    tot[0] += 1;
    tot[1] += st->iterations;
    tot[2] += st->converged ? 0 : 1;
    tot[3] += st->ap_evals;
    tot[4] |= st->status;
    tot[5] = tot[5] > (long long)st->bar_seen ? tot[5] : (long long)st->bar_seen;
    tot[6] = end.grid;
  }
}

__global__ void k_reset_state(DevState *st, int *list_max) {
  const int t = threadIdx.x;
  for (int k = t; k < kMaxIters; k += blockDim.x) {
    st->res_sum[k] = 0ull;
    st->res_cnt[k] = 0u;
  }
  for (int k = t; k < kListMaxSlots; k += blockDim.x) {
    list_max[k * kListMaxStride] = 0;
    list_max[k * kListMaxStride + 1] = 0;
  }
  if (t == 0) {
    st->vmax_bits = 0u;
    st->iterations = 0;
    st->ap_evals = 0;
    st->converged = 0;
    st->max_ff = 0;
    st->max_fs = 0;
    st->status = 0;
    st->done = 0;
    st->step_tag += 1;
    st->iter = 0;
    st->final_src = 0;
    st->residual = 0.0;
    st->bar_count = 0u;
    st->bar_gen = 0u;
    st->bar_seen = 0u;
  }
}

// ---- persistent cooperative Jacobi loop -----------------------------------
// Grid-wide barrier with a bounded wait: a wait that exceeds the spin limit
// raises ST_BARRIER and every block leaves the loop, so the kernel always
// ends.
constexpr long kBarrierSpinLimit = 1L << 19;

__device__ __forceinline__ bool grid_barrier(DevState *st) {
  __shared__ int s_abort;
  // every wave's stores have completed at the workgroup barrier; one thread
  // per block then publishes them at agent scope and waits
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nblk = gridDim.x;
    const unsigned gen =
        __hip_atomic_load(&st->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned arrived =
        __hip_atomic_fetch_add(&st->bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblk - 1) {
      __hip_atomic_store(&st->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&st->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&st->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if ((++spins & 255) == 0 &&
            (__hip_atomic_load(&st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
             ST_BARRIER))
          break;
        if (spins > kBarrierSpinLimit) {
          atomicMax(&st->bar_seen,
                    __hip_atomic_load(&st->bar_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          atomicOr(&st->status, (int)ST_BARRIER);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_abort = __hip_atomic_load(&st->status, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) &
              ST_BARRIER;
  }
  __syncthreads();
  return s_abort == 0;
}

template <bool kLds>
__global__ void k_jacobi_persistent(Bufs b, SphConst c, int src, int min_it, int max_it,
                                    int force_it, double eta) {
  const Tabs T = stage_tabs<kLds>(b, c);
  const float dt = (float)b.st->dt;
  const float t3 = c.mass * (dt * dt);
  // each block owns a contiguous range of particles (XCD-aware order)
  const int per_block = ((b.n + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
  const int lo = xcd_tile() * per_block;
  const int hi = min(b.n, lo + per_block);
  const int stride = blockDim.x;
  const int first = lo + threadIdx.x;
  int cur = src;
  int k = 0;
  double e = 0.0;
  bool ok = true;
  if (b.scenes) {
    // batched scenes: every block takes the same stop decisions from the
    // scenes' residual sums and keeps them in LDS; block 0 records them
    __shared__ unsigned char s_done[kMaxScenes];
    __shared__ unsigned char s_final[kMaxScenes];
    for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x) s_done[sc] = 0;
    __syncthreads();
    for (;; ++k) {
      for (int i0 = lo; i0 < hi; i0 += stride) {
        const int i = i0 + threadIdx.x;
        if (i < hi && !s_done[b.sid[i]]) j1_particle(b, c, T, i, cur, t3);
      }
      if (!(ok = grid_barrier(b.st))) break;
      for (int i0 = lo; i0 < hi; i0 += stride) {
        const int i = i0 + threadIdx.x;
        const int sc = i < hi ? b.sid[i] : -1;
        const bool act = sc >= 0 && !s_done[sc];
        unsigned long long rs = 0;
        unsigned rc = 0;
        if (act) j2_particle(b, c, T, i, cur, rs, rc);
        scene_residual_add(b, k, act ? sc : -1, rs, rc);
      }
      if (!(ok = grid_barrier(b.st))) break;
      for (int sc = threadIdx.x; sc < b.nscenes; sc += blockDim.x)
        if (!s_done[sc]) s_final[sc] = (unsigned char)cur;
      __syncthreads();
      if (!batch_decide(b, k, cur, min_it, max_it, force_it, eta, s_done, blockIdx.x == 0)) break;
      cur ^= 1;
    }
    // leave every scene's final iterate in dp[0]
    if (ok)
      for (int i = first; i < hi; i += stride)
        if (s_final[b.sid[i]] == 1) b.dp[0][i] = b.dp[1][i];
    if (blockIdx.x == 0) {
      double emax = 0.0;
      int conv = 1;
      for (int sc = 0; sc < b.nscenes; ++sc) {
        const double es = b.sc_res[sc];
        emax = es > emax ? es : emax;
        conv &= es <= eta;
      }
      if (threadIdx.x == 0) {
        b.st->iterations = k + 1;
        b.st->ap_evals += k + 1;
        b.st->residual = emax;
        b.st->converged = conv;
      }
    }
    return;
  }
  for (;; ++k) {
    for (int i = first; i < hi; i += stride) j1_particle(b, c, T, i, cur, t3);
    if (!(ok = grid_barrier(b.st))) break;
    unsigned long long rs = 0;
    unsigned rc = 0;
    for (int i = first; i < hi; i += stride) j2_particle(b, c, T, i, cur, rs, rc);
    block_residual_add(b.st, k, rs, rc);
    if (!(ok = grid_barrier(b.st))) break;
    const unsigned long long S =
        __hip_atomic_load(&b.st->res_sum[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned Cn =
        __hip_atomic_load(&b.st->res_cnt[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (!jacobi_continue(k, S, Cn, min_it, max_it, force_it, eta, &e)) break;
    cur ^= 1;
  }
  // leave the final iterate in dp[0]
  if (ok && cur == 1)
    for (int i = first; i < hi; i += stride) b.dp[0][i] = b.dp[1][i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.st->iterations = k + 1;
    b.st->ap_evals += k + 1;
    b.st->residual = e;
    b.st->converged = e <= eta;
  }
}

// ---- scatter a field back to original order ------------------------------
__global__ void k_scatter(const unsigned *__restrict__ src, int stride, int ncomp, int off,
                          const int *__restrict__ orig, int n, unsigned *__restrict__ dst) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const long o = orig[f];
  for (int k = 0; k < ncomp; ++k) dst[o * ncomp + k] = src[(long)f * stride + off + k];
}

// ---- launchers ---------------------------------------------------------------
static inline unsigned nblocks(int n, int block) { return (unsigned)((n + block - 1) / block); }
static inline size_t tab_bytes(const Bufs &b, const SphConst &c) {
  return b.lds_tables ? sizeof(float) * (size_t)(c.nw + c.ng) : 0;
}

// Dynamic LDS of a J2 launch: the tables, raised so that at most 4 waves per
// SIMD (16 per CU) are resident. J2's neighbour gathers hit L1 more often with
// fewer waves sharing it: measured at 1M particles, 512-thread blocks, 92 us
// at 2 blocks per CU against 96 us at 3 (the register limit) and 115 us at 1.
static inline size_t j2_lds_bytes(const Bufs &b, const SphConst &c, int block) {
  const size_t tab = tab_bytes(b, c);
  if (b.lds_per_cu <= 0) return tab;
  const size_t cap = (size_t)b.lds_per_cu * (size_t)(block / 64) / 16;
  const size_t want = cap > 1024 ? cap - 1024 : 0;  // below the share, whatever the granule
  return tab > want ? tab : want;
}

#define DISPATCH_LDS(kern, grid, block, shm, stream, ...)                    \
  do {                                                                      \
    if (b.lds_tables)                                                       \
      kern<true><<<(grid), (block), (shm), (stream)>>>(__VA_ARGS__);        \
    else                                                                    \
      kern<false><<<(grid), (block), (shm), (stream)>>>(__VA_ARGS__);       \
  } while (0)

void launch_boundary_psi(const Bufs &b, const SphConst &c, hipStream_t s) {
  if (b.nb == 0) return;
  DISPATCH_LDS(k_boundary_psi, nblocks(b.nb, 256), 256, tab_bytes(b, c), s, b, c);
}

// ---- self-test of the rounding helpers -------------------------------------
// bad[0]: float32 bit patterns b (all 2^32) where rcp_rn(b) != 1.0f / b.
// bad[1]: non-negative b up to +inf where sqrt_rn(b) != sqrtf(b).
// bad[2]: pairs where div_by(a, b, rcp_rn(b)) != a / b, over every significand
//   of b at 64 exponents in [-28, 35], each with a pseudo-random normal a of
//   exponent in [-27, 27] (536,870,912 pairs; quotients are normal).
// bad[3]: non-negative b up to +inf where sqrt_rn_pair(b) differs from sqrtf(b)
//   (b >= 2^-96) or is not below 2^-47 (b < 2^-96).
__global__ void k_selftest_arith(unsigned long long *bad) {
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  unsigned nr = 0, ns = 0, nd = 0, np = 0;
  for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       u < (1ull << 32); u += stride) {
    const unsigned bits = (unsigned)u;
    const float b = __uint_as_float(bits);
    nr += __float_as_uint(rcp_rn(b)) != __float_as_uint(1.0f / b);
    if (bits <= 0x7f800000u) {
      ns += __float_as_uint(sqrt_rn(b)) != __float_as_uint(sqrtf(b));
      const float sp = sqrt_rn_pair(b);
      np += bits >= 0x0f800000u ? __float_as_uint(sp) != __float_as_uint(sqrtf(b)) : !(sp < 0x1.0p-47f);
    }
    if (bits < (1u << 29)) {
      const unsigned e = 99u + (bits >> 23);  // 64 exponents
      const float bb = __uint_as_float((e << 23) | (bits & 0x7fffffu));
      unsigned h = bits * 0x9e3779b9u;
      h ^= h >> 16;
      h *= 0x85ebca6bu;
      h ^= h >> 13;
      const float a = __uint_as_float(((100u + (h >> 23) % 55u) << 23) | (h & 0x7fffffu));
      nd += __float_as_uint(div_by(a, bb, rcp_rn(bb))) != __float_as_uint(a / bb);
    }
  }
  const unsigned long long r = wave_sum_u64(nr), q = wave_sum_u64(ns), d = wave_sum_u64(nd),
                           e = wave_sum_u64(np);
  if ((threadIdx.x & 63) == 0) {
    if (r) atomicAdd(&bad[0], r);
    if (q) atomicAdd(&bad[1], q);
    if (d) atomicAdd(&bad[2], d);
    if (e) atomicAdd(&bad[3], e);
  }
}

void launch_selftest_arith(unsigned long long *bad, hipStream_t s) {
  k_selftest_arith<<<4096, 256, 0, s>>>(bad);
}

void launch_mark_bnear(const Bufs &b, const SphConst &c, unsigned *bnear, hipStream_t s) {
  if (b.nb > 0) k_mark_bnear<<<nblocks(b.nb, 256), 256, 0, s>>>(b, c, bnear);
}

// k_lists runs in blocks of kListsBlock threads whatever block the other
// passes use: its LDS ring (8 KB per wave) and 100 registers then admit 4
// waves per SIMD. It is bound by the texture data path (about two cycles per
// cache line a wave's gather touches; its lanes sit in a dozen or more query
// cells, so a gather touches ~19 lines), not by its VALU work or occupancy:
// 4 or 5 waves per SIMD, one or two batches in flight all measured 422-425 us
// at 1M particles.
constexpr int kListsBlock = 128;

void launch_lists(const Bufs &b, const SphConst &c, int /*other passes' block*/, hipStream_t s) {
  if (b.n == 0) return;
  const int block = kListsBlock;
  const size_t ring = sizeof(int) * (size_t)(block / 64) * kListRing * 64;
  k_lists<<<nblocks(b.n, block), block, ring, s>>>(b, c);
}

void launch_density(const Bufs &b, const SphConst &c, bool fuse_vadv, int block, hipStream_t s) {
  if (b.n == 0) return;
  DISPATCH_LDS(k_density, nblocks(b.n, block), block, tab_bytes(b, c), s, b, c, (int)fuse_vadv);
}

void launch_vadv(const Bufs &b, const SphConst &c, int block, hipStream_t s) {
  if (b.n == 0) return;
  DISPATCH_LDS(k_vadv, nblocks(b.n, block), block, tab_bytes(b, c), s, b, c);
}

void launch_assemble(const Bufs &b, const SphConst &c, int start_mode, int block, hipStream_t s) {
  if (b.n == 0) return;
  DISPATCH_LDS(k_assemble, nblocks(b.n, block), block, tab_bytes(b, c), s, b, c, start_mode);
}

void launch_j1(const Bufs &b, const SphConst &c, int src, int block, hipStream_t s) {
  if (b.n == 0) return;
  DISPATCH_LDS(k_j1, nblocks(b.n, block), block, tab_bytes(b, c), s, b, c, src);
}

void launch_j2(const Bufs &b, const SphConst &c, int src, int slot, int block, hipStream_t s) {
  if (b.n == 0) return;
  DISPATCH_LDS(k_j2, nblocks(b.n, block), block, j2_lds_bytes(b, c, block), s, b, c, src, slot);
  if (slot >= 0)
    k_residual_reduce<<<1, 1024, 0, s>>>(b, (int)nblocks(b.n, block), slot);
}

void launch_jacobi_chunk(const Bufs &b, const SphConst &c, const wsph_params &P, int src0,
                         int k0, int iters, int block, hipStream_t s, hipEvent_t *ev3) {
  if (b.n == 0) return;
  const LoopCtl L = loop_ctl(P);
  for (int k = k0; k < k0 + iters; ++k) {
    const bool ev = ev3 && ev3[0] && k == k0;
    if (ev) (void)hipEventRecord(ev3[0], s);
    DISPATCH_LDS(k_j1_dev, nblocks(b.n, block), block, tab_bytes(b, c), s, b, c, src0, k, L);
    if (ev && ev3[1]) (void)hipEventRecord(ev3[1], s);
    DISPATCH_LDS(k_j2_dev, nblocks(b.n, block), block, j2_lds_bytes(b, c, block), s, b, c, src0, k);
    if (ev) (void)hipEventRecord(ev3[2], s);
    if (b.scenes)
      k_batch_stop<<<1, 1024, 0, s>>>(b, src0, P.min_iters, P.max_iters, P.force_iters, P.eta);
  }
}

void launch_jacobi_last_test(const Bufs &b, const wsph_params &P, int src0, int k, hipStream_t s) {
  if (b.n == 0 || b.scenes || k < 0) return;
  k_jacobi_last_test<<<1, 1024, 0, s>>>(b, (int)nblocks(b.n, P.block), src0, k, loop_ctl(P));
}

void launch_batch_finalize(const Bufs &b, hipStream_t s) {
  if (b.n == 0 || !b.scenes) return;
  k_batch_finalize<<<nblocks(b.n, 256), 256, 0, s>>>(b);
}

void launch_integrate(const Bufs &b, const SphConst &c, int src, int start_mode, int block,
                      const StepEnd &end, hipStream_t s) {
  if (b.n > 0) k_integrate<<<nblocks(b.n, block), block, 0, s>>>(b, c, src);
  k_vmax_reduce<<<1, 1024, 0, s>>>(b, b.n > 0 ? (int)nblocks(b.n, block) : 0, start_mode, end);
}

void launch_reset_state(const Bufs &b, hipStream_t s) {
  k_reset_state<<<1, 128, 0, s>>>(b.st, b.list_max);
  if (b.scenes) k_batch_reset<<<64, 256, 0, s>>>(b, kMaxIters);
}

// Blocks of k_jacobi_persistent that fit on one CU at once, from the kernel's
// attributes (VGPRs, static LDS -- the barrier's and the residual's
// __shared__ words -- plus the dynamic table LDS) and the device's limits;
// *total = that figure times the number of CUs.
int persistent_blocks_per_cu(const Bufs &b, const SphConst &c, int block, int *total) {
  int dev = 0, ncu = 0, lds_cu = 0, threads_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
          hipSuccess ||
      hipDeviceGetAttribute(&threads_cu, hipDeviceAttributeMaxThreadsPerMultiProcessor, dev) !=
          hipSuccess)
    return -1;
  hipFuncAttributes fa;
  const hipError_t e = b.lds_tables
                           ? hipFuncGetAttributes(&fa, (const void *)k_jacobi_persistent<true>)
                           : hipFuncGetAttributes(&fa, (const void *)k_jacobi_persistent<false>);
  if (e != hipSuccess) return -1;
  const int lds = (int)(fa.sharedSizeBytes + tab_bytes(b, c));
  // the runtime reports no SGPR count: 0 takes the largest a kernel can hold
  const int per = wsph_blocks_per_cu(block, fa.numRegs, 0, lds, lds_cu, threads_cu / 64 / 4);
  if (total) *total = per * ncu;
  return per;
}

hipError_t launch_jacobi_persistent(const Bufs &b, const SphConst &c, const wsph_params &P,
                                    int src, int grid, int block, hipStream_t s) {
  Bufs bb = b;
  SphConst cc = c;
  int min_it = P.min_iters, max_it = P.max_iters, force_it = P.force_iters;
  double eta = P.eta;
  void *args[] = {&bb, &cc, &src, &min_it, &max_it, &force_it, &eta};
  const unsigned shm = (unsigned)tab_bytes(b, c);
  if (b.lds_tables)
    return hipLaunchCooperativeKernel((const void *)k_jacobi_persistent<true>, dim3(grid),
                                      dim3(block), args, shm, s);
  return hipLaunchCooperativeKernel((const void *)k_jacobi_persistent<false>, dim3(grid),
                                    dim3(block), args, shm, s);
}

void launch_scatter(const void *src, int stride_floats, int ncomp, int src_off, const int *orig,
                    int n, void *dst, hipStream_t s) {
  if (n == 0) return;
  k_scatter<<<nblocks(n, 256), 256, 0, s>>>((const unsigned *)src, stride_floats, ncomp, src_off,
                                            orig, n, (unsigned *)dst);
}

}  // namespace wsph
