// mesh.hip -- lattice points inside a closed triangle mesh.
//
// The reference samples its model boundaries (model.py:14-47) by keeping the
// points of an np.arange lattice over the scaled mesh's bounds that
// trimesh (4.6.10) reports inside (Trimesh.contains: ray casting with an
// intersection-parity count). This is the same test, made deterministic:
//
// * rays run along +z, one per lattice column (x_i, y_j); a column crosses
//   triangle t when (x_i, y_j) lies inside t's xy projection;
// * ties -- a column through an edge or a vertex of the projection -- follow
//   the rasterisation top-left rule on edge functions that the two triangles
//   of a shared edge compute with exactly opposite values (endpoints taken in
//   lexicographic order). For a watertight mesh a column through a shared
//   edge then crosses exactly one of its two triangles (both or neither at a
//   silhouette fold, where the ray grazes), and a column through a vertex
//   crosses one triangle of the fan around it;
// * a point is inside when an odd number of its column's crossings lie
//   strictly above it.
//
// Everything is double precision +, -, *, / in a fixed order (the library is
// built without fused multiply-add), so any device gives the same answer.
// Work: the columns are grouped in tiles of 16 x 16; each triangle is binned
// into the tiles whose columns its closed xy box covers (binary search on
// the ascending axes, a counting pass and a fill pass). A block takes one
// tile, stages its triangles through LDS and each column tests them; a count
// pass sizes the per-column crossing lists, a second pass fills them, a third
// sorts each list and classifies the column's points with one sweep.
#include "kernels.h"

namespace wsph {

constexpr int kMeshBlock = 256;

// Edge function of P->Q at R with the endpoints in lexicographic order.
__device__ __forceinline__ double edge_fn(double px, double py, double qx, double qy, double rx,
                                          double ry) {
  const bool sw = qx < px || (qx == px && qy < py);
  if (sw) {
    double t = px;
    px = qx;
    qx = t;
    t = py;
    py = qy;
    qy = t;
  }
  const double e = (qx - px) * (ry - py) - (qy - py) * (rx - px);
  return sw ? -e : e;
}

// Top-left rule for the edge P->Q of a counter-clockwise triangle (y up):
// left edges run downwards, top edges are horizontal and run towards -x.
__device__ __forceinline__ bool top_left(double px, double py, double qx, double qy) {
  return qy < py || (qy == py && qx < px);
}

// Height of the crossing of column (rx, ry) with t, when it crosses.
__device__ __forceinline__ bool crossing(const MeshTri &t, double rx, double ry, double &z) {
  if (t.orient == 0 || rx < t.lo[0] || rx > t.hi[0] || ry < t.lo[1] || ry > t.hi[1]) return false;
  double w[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int p = e == 2 ? 0 : e + 1, q = e == 0 ? 2 : e - 1;  // edge opposite vertex e
    w[e] = edge_fn(t.x[p], t.y[p], t.x[q], t.y[q], rx, ry);
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int p = e == 2 ? 0 : e + 1, q = e == 0 ? 2 : e - 1;
    const double we = t.orient > 0 ? w[e] : -w[e];
    const bool tl = t.orient > 0 ? top_left(t.x[p], t.y[p], t.x[q], t.y[q])
                                 : top_left(t.x[q], t.y[q], t.x[p], t.y[p]);
    if (!(we > 0.0 || (we == 0.0 && tl))) return false;
  }
  const double s = (w[0] + w[1]) + w[2];
  z = ((w[0] * t.z[0] + w[1] * t.z[1]) + w[2] * t.z[2]) / s;
  return true;
}

// Column tiles: 16 x 16 columns (i, j) per block, tile (i / 16, j / 16).
constexpr int kMeshTile = 16;
static_assert(kMeshTile * kMeshTile == kMeshBlock, "one thread per column of the tile");
// triangles per work item of the column passes (enough blocks to fill the
// device when a few tiles hold most of the triangles)
constexpr int kMeshChunk = 32;
static_assert(kMeshChunk <= kMeshBlock, "one triangle per thread when staging");

// first index with a[index] >= v, last index with a[index] <= v (a ascending)
__device__ __forceinline__ int first_ge(const double *__restrict__ a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int last_le(const double *__restrict__ a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

// Triangles from the vertex/index arrays, with the range of column tiles
// their closed xy box covers (the columns whose own box test can pass), and
// the number of triangles per tile.
__global__ void k_mesh_setup(const double *__restrict__ verts, int nv, const int *__restrict__ tris,
                             int nt, const double *__restrict__ ax, int nx,
                             const double *__restrict__ ay, int ny, MeshTri *__restrict__ out,
                             unsigned *__restrict__ tile_cnt, int *__restrict__ bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  MeshTri m;
  int idx[3];
  bool ok = true;
  for (int k = 0; k < 3; ++k) {
    idx[k] = tris[3 * (long)t + k];
    ok = ok && idx[k] >= 0 && idx[k] < nv;
  }
  if (!ok) {
    atomicOr(bad, 1);
    idx[0] = idx[1] = idx[2] = 0;
  }
  for (int k = 0; k < 3; ++k) {
    m.x[k] = verts[3 * (long)idx[k]];
    m.y[k] = verts[3 * (long)idx[k] + 1];
    m.z[k] = verts[3 * (long)idx[k] + 2];
  }
  m.lo[0] = fmin(fmin(m.x[0], m.x[1]), m.x[2]);
  m.hi[0] = fmax(fmax(m.x[0], m.x[1]), m.x[2]);
  m.lo[1] = fmin(fmin(m.y[0], m.y[1]), m.y[2]);
  m.hi[1] = fmax(fmax(m.y[0], m.y[1]), m.y[2]);
  // orientation of the projection (vertex 2 against the edge 0->1)
  const double a = edge_fn(m.x[0], m.y[0], m.x[1], m.y[1], m.x[2], m.y[2]);
  m.orient = !ok ? 0 : a > 0.0 ? 1 : a < 0.0 ? -1 : 0;
  const int i0 = first_ge(ax, nx, m.lo[0]), i1 = last_le(ax, nx, m.hi[0]);
  const int j0 = first_ge(ay, ny, m.lo[1]), j1 = last_le(ay, ny, m.hi[1]);
  if (m.orient == 0 || i0 > i1 || j0 > j1) {
    m.tiles[0] = 0;
    m.tiles[1] = -1;  // no tile
    m.tiles[2] = 0;
    m.tiles[3] = -1;
  } else {
    m.tiles[0] = i0 / kMeshTile;
    m.tiles[1] = i1 / kMeshTile;
    m.tiles[2] = j0 / kMeshTile;
    m.tiles[3] = j1 / kMeshTile;
  }
  out[t] = m;
  const int tiles_y = (ny + kMeshTile - 1) / kMeshTile;
  for (int ti = m.tiles[0]; ti <= m.tiles[1]; ++ti)
    for (int tj = m.tiles[2]; tj <= m.tiles[3]; ++tj) atomicAdd(&tile_cnt[ti * tiles_y + tj], 1u);
}

// Triangle ids per tile (order within a tile is irrelevant).
__global__ void k_mesh_bin(const MeshTri *__restrict__ tris, int nt, int ny,
                           const unsigned *__restrict__ tile_off, unsigned *__restrict__ tile_pos,
                           int *__restrict__ ids) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int tiles_y = (ny + kMeshTile - 1) / kMeshTile;
  const int *r = tris[t].tiles;
  for (int ti = r[0]; ti <= r[1]; ++ti)
    for (int tj = r[2]; tj <= r[3]; ++tj) {
      const int k = ti * tiles_y + tj;
      ids[tile_off[k] + atomicAdd(&tile_pos[k], 1u)] = t;
    }
}

// One block per work item (a tile and a chunk of at most kMeshChunk of its
// triangles), one thread per column (i, j) of the tile, c = i * ny + j.
// kFill == false adds the chunk's crossings to cnt[c] (zeroed by the
// caller); kFill == true reserves that many slots of the column's list with
// the cursor fill[c] (zeroed) and writes their heights. Counts are sums and
// the heights are sorted afterwards, so the result does not depend on the
// order in which the chunks run.
template <bool kFill>
__global__ __launch_bounds__(kMeshBlock) void k_mesh_columns(
    const MeshTri *__restrict__ tris, const int *__restrict__ ids,
    const int4 *__restrict__ items, const double *__restrict__ ax, int nx,
    const double *__restrict__ ay, int ny, unsigned *__restrict__ cnt,
    const unsigned *__restrict__ off, unsigned *__restrict__ fill, double *__restrict__ zbuf) {
  __shared__ MeshTri tile[kMeshChunk];
  const int4 it = items[blockIdx.x];  // tile, first id, number of ids
  const int tiles_y = (ny + kMeshTile - 1) / kMeshTile;
  const int i = (it.x / tiles_y) * kMeshTile + (int)threadIdx.x / kMeshTile;
  const int j = (it.x % tiles_y) * kMeshTile + (int)threadIdx.x % kMeshTile;
  const bool live = i < nx && j < ny;
  const long c = (long)i * ny + j;
  const double rx = live ? ax[i] : 0.0, ry = live ? ay[j] : 0.0;
  if ((int)threadIdx.x < it.z) tile[threadIdx.x] = tris[ids[it.y + threadIdx.x]];
  __syncthreads();
  if (!live) return;
  unsigned m = 0;
  double z;
  for (int k = 0; k < it.z; ++k) m += crossing(tile[k], rx, ry, z) ? 1u : 0u;
  if (m == 0) return;
  if (!kFill) {
    atomicAdd(&cnt[c], m);
    return;
  }
  double *out = zbuf + off[c] + atomicAdd(&fill[c], m);
  unsigned w = 0;
  for (int k = 0; k < it.z; ++k)
    if (crossing(tile[k], rx, ry, z)) out[w++] = z;
}

// Axes strictly ascending (the tile ranges and the sweep rely on it).
__global__ void k_mesh_check_axes(const double *__restrict__ a, int n, int *__restrict__ bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k + 1 < n && !(a[k] < a[k + 1])) atomicOr(bad, 2);
}

// Sort each column's crossing heights (ascending, insertion sort: columns
// cross few triangles).
__global__ void k_mesh_sort(const unsigned *__restrict__ off, const unsigned *__restrict__ cnt,
                            double *__restrict__ zbuf, long ncol) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double *z = zbuf + off[c];
  const int m = (int)cnt[c];
  for (int a = 1; a < m; ++a) {
    const double v = z[a];
    int b = a - 1;
    while (b >= 0 && z[b] > v) {
      z[b + 1] = z[b];
      --b;
    }
    z[b + 1] = v;
  }
}

// One thread per point p = c * nz + k (consecutive threads store consecutive
// bytes): inside when an odd number of the column's sorted crossings lie
// strictly above az[k].
__global__ void k_mesh_mark(const unsigned *__restrict__ off, const unsigned *__restrict__ cnt,
                            const double *__restrict__ zbuf, long npts,
                            const double *__restrict__ az, int nz,
                            unsigned char *__restrict__ inside) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long c = p / nz;
  const double q = az[p - c * nz];
  const double *z = zbuf + off[c];
  int lo = 0, hi = (int)cnt[c];  // first crossing above q
  const int m = hi;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (z[mid] <= q)
      lo = mid + 1;
    else
      hi = mid;
  }
  inside[p] = (unsigned char)((m - lo) & 1);
}

// Exclusive prefix sum of n counts in one block: each thread sums one
// contiguous chunk, the block scans the chunk sums (Hillis-Steele, 64-bit),
// and each thread writes its chunk's offsets. With per > 0 entry k counts
// ceil(in[k] / per) (the work items of a tile) instead of in[k]. *total
// receives the 64-bit sum; offsets are stored in 32 bits, so a caller uses
// them only once it has checked *total.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan_excl(const unsigned *__restrict__ in, long n,
                                                          unsigned per, unsigned *__restrict__ out,
                                                          unsigned long long *__restrict__ total) {
  __shared__ unsigned long long s[kScanBlock];
  const long chunk = (n + kScanBlock - 1) / kScanBlock;
  const long b = (long)threadIdx.x * chunk;
  const long e = b + chunk < n ? b + chunk : n;
  auto val = [&](long k) -> unsigned long long {
    return per ? ((unsigned long long)in[k] + per - 1) / per : in[k];
  };
  unsigned long long sum = 0;
  for (long k = b; k < e; ++k) sum += val(k);
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < kScanBlock; o <<= 1) {
    const unsigned long long v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned long long run = s[threadIdx.x] - sum;
  for (long k = b; k < e; ++k) {
    out[k] = (unsigned)run;
    run += val(k);
  }
  if (threadIdx.x == kScanBlock - 1) *total = s[kScanBlock - 1];
}

// Work items of the column passes, tile by tile (one thread per tile): the
// tile's triangles in chunks of at most `chunk`, {tile, first id, count, 0}
// at item_off[tile] onwards.
__global__ void k_mesh_items(const unsigned *__restrict__ tile_cnt,
                             const unsigned *__restrict__ tile_off,
                             const unsigned *__restrict__ item_off, long ntiles, int chunk,
                             int4 *__restrict__ items) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const unsigned c = tile_cnt[t];
  unsigned q = item_off[t];
  for (unsigned f = 0; f < c; f += (unsigned)chunk, ++q) {
    const unsigned left = c - f;
    items[q] = make_int4((int)t, (int)(tile_off[t] + f),
                         (int)(left < (unsigned)chunk ? left : (unsigned)chunk), 0);
  }
}

void launch_scan_excl(const unsigned *in, long n, unsigned per, unsigned *out,
                      unsigned long long *total, hipStream_t st) {
  k_scan_excl<<<1, kScanBlock, 0, st>>>(in, n, per, out, total);
}

void launch_mesh_items(const unsigned *tile_cnt, const unsigned *tile_off,
                       const unsigned *item_off, long ntiles, int4 *items, hipStream_t st) {
  if (ntiles > 0)
    k_mesh_items<<<(unsigned)((ntiles + 255) / 256), 256, 0, st>>>(tile_cnt, tile_off, item_off,
                                                                   ntiles, kMeshChunk, items);
}

void launch_mesh_check_axes(const double *a, int n, int *bad, hipStream_t st) {
  if (n > 1) k_mesh_check_axes<<<(n + 255) / 256, 256, 0, st>>>(a, n, bad);
}

long mesh_tiles(int nx, int ny) {
  return ((long)nx + kMeshTile - 1) / kMeshTile * (((long)ny + kMeshTile - 1) / kMeshTile);
}

void launch_mesh_setup(const double *verts, int nv, const int *tris, int nt, const double *ax,
                       int nx, const double *ay, int ny, MeshTri *out, unsigned *tile_cnt,
                       int *bad, hipStream_t st) {
  if (nt > 0)
    k_mesh_setup<<<(nt + 255) / 256, 256, 0, st>>>(verts, nv, tris, nt, ax, nx, ay, ny, out,
                                                   tile_cnt, bad);
}

void launch_mesh_bin(const MeshTri *tris, int nt, int ny, const unsigned *tile_off,
                     unsigned *tile_pos, int *ids, hipStream_t st) {
  if (nt > 0) k_mesh_bin<<<(nt + 255) / 256, 256, 0, st>>>(tris, nt, ny, tile_off, tile_pos, ids);
}

int mesh_chunk() { return kMeshChunk; }

void launch_mesh_columns(const MeshTri *tris, const int *ids, const int4 *items, int nitems,
                         const double *ax, int nx, const double *ay, int ny, unsigned *cnt,
                         const unsigned *off, unsigned *fill, double *zbuf, hipStream_t st) {
  if (nitems == 0) return;
  if (zbuf)
    k_mesh_columns<true><<<(unsigned)nitems, kMeshBlock, 0, st>>>(tris, ids, items, ax, nx, ay, ny,
                                                                  cnt, off, fill, zbuf);
  else
    k_mesh_columns<false><<<(unsigned)nitems, kMeshBlock, 0, st>>>(tris, ids, items, ax, nx, ay,
                                                                   ny, cnt, off, fill, zbuf);
}

void launch_mesh_classify(const unsigned *off, const unsigned *cnt, double *zbuf, long ncol,
                          const double *az, int nz, unsigned char *inside, hipStream_t st) {
  if (ncol == 0 || nz == 0) return;
  k_mesh_sort<<<(unsigned)((ncol + 255) / 256), 256, 0, st>>>(off, cnt, zbuf, ncol);
  const long npts = ncol * nz;
  k_mesh_mark<<<(unsigned)((npts + 255) / 256), 256, 0, st>>>(off, cnt, zbuf, npts, az, nz, inside);
}

}  // namespace wsph
