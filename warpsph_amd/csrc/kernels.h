// kernels.h -- internal interface between the C-ABI layer and the kernels.
#pragma once

#include "common.h"

namespace wsph {

// ---- hash grid build (grid.hip) ----------------------------------------
struct GridArgs {
  int n;
  long ncell;
  float inv;
  int dims[3];
  int layout;
  const float4 *pos;   // current order
  const int *orig;     // original index of each current slot
  int *slot, *rank, *tmp, *perm;
  unsigned *count, *start;
  unsigned **scan_bufs;
  // batched scenes (null: one scene with `dims`)
  const int *sid;            // scene of each point, current order
  const SceneGrid *scenes;
  int boundary;              // 1: use the scenes' boundary grids
};

int scan_levels_needed(long n, long *sizes, int max_levels);
void exclusive_scan(const unsigned *in, unsigned *out, long n, unsigned **level_bufs,
                    hipStream_t st);
void sort_into_cells(const GridArgs &g, hipStream_t st);
void gather_fluid(const int *perm, int n, const float4 *pos, const float4 *vel, const float *pprev,
                  const int *orig, float4 *pos_o, float4 *vel_o, float *pprev_o, int *orig_o,
                  hipStream_t st);
void fill_u32(unsigned *p, unsigned v, long n, hipStream_t st);
void gather_i32(const int *perm, int n, const int *src, int *dst, hipStream_t st);
void linear_keys(const float *x, int n, float inv, const int *dims, int *keys, hipStream_t st);

// ---- SPH passes (sph.hip) ------------------------------------------------
// Every per-particle array is in cell-sorted order.
struct Bufs {
  int n, nb;
  // fluid
  float4 *pos;      // x, y, z, rho
  float4 *vel;      // v
  float4 *vadv;     // v_adv
  float4 *dp[2];    // d_ii.xyz, p   (double buffer of the Jacobi iterate)
  float4 *sap;      // s = sum_j d_ij p_j (xyz), first part of A p (w)
  float4 *aux;      // rho_adv, a_ii
  float *ap;        // A p
  float *pprev;     // pressure carried to the next step's start value
  float *pfinal;    // final pressure of the last step
  unsigned *vmax_part;  // per-block max |v| of k_integrate
  double *ke_part;      // per-block kinetic energy of k_integrate
  double *y_part;       // per-block sum of heights
  unsigned *ymax_part;  // per-block max height (ordered bits)
  StepRecord *history;  // kHistory entries
  unsigned long long *res_part;  // per-block residual sums of J2
  // largest neighbour counts of the list pass, spread over kListMaxSlots
  // cache lines (fluid at slot * 32, boundary at slot * 32 + 1); folded into
  // DevState::max_ff / max_fs at the step's end
  int *list_max;
  unsigned *cnt_part;            // per-block residual counts of J2
  const float *pstart;  // caller-supplied start pressure, sorted order (or null)
  // pressure predictor (null / 0 when unused)
  float4 *seed;         // per-particle self features 0..3, written by k_assemble
  float *feat;          // feature rows [n][8] (pair logging)
  float *target;        // regression targets [n] (pair logging)
  double *err_part;     // per-block (net, carry) squared start errors of k_integrate
  int arbiter;            // learned start: compare both start values after the solve
  // batched scenes (scenes == null: a single scene)
  const SceneGrid *scenes;
  const int *sid;       // fluid scene id, sorted order
  const int *bsid;      // boundary scene id, sorted order
  int nscenes;
  unsigned long long *sc_sum;  // [kMaxIters][nscenes] residual sums per iteration
  unsigned *sc_cnt;            // [kMaxIters][nscenes] residual counts
  int *sc_done;         // scene finished its Jacobi loop this step
  int *sc_iters;        // iterations of the scene this step
  int *sc_final;        // dp buffer holding the scene's final iterate
  double *sc_res;       // residual at the scene's exit
  // fluid grid
  const unsigned *fstart, *fcount;
  // neighbour lists, tiled by wave: entry k of i at list[list_base(i, cap) + k * 64]
  int *nbr_ff, *cnt_ff, *nbr_fs, *cnt_fs;
  int cap_ff, cap_fs;
  // boundary (sorted by its own grid)
  float4 *bpos;     // x, y, z, psi
  float4 *bvel;
  const unsigned *bstart, *bcount;
  // per boundary slot: nonzero when a boundary particle's cell lies within 2
  // cells on every axis (null: not known, every query searches)
  const unsigned *bnear;
  int *done_host;  // pinned host memory: the stop test stores the finished step's tag
  unsigned long long *prog_host;  // pinned: (step tag << 32) | iterations tested
  // tables
  const float *wt, *gt;
  bool lds_tables;
  int lds_per_cu;  // bytes of LDS per compute unit (host side: J2 occupancy cap)
  DevState *st;
};

void launch_boundary_psi(const Bufs &b, const SphConst &c, hipStream_t s);
void launch_lists(const Bufs &b, const SphConst &c, int block, hipStream_t s);
// mark the boundary slots near boundary particles (bnear zeroed by the caller)
void launch_mark_bnear(const Bufs &b, const SphConst &c, unsigned *bnear, hipStream_t s);
void launch_selftest_arith(unsigned long long *bad, hipStream_t s);
void launch_density(const Bufs &b, const SphConst &c, bool fuse_vadv, int block, hipStream_t s);
void launch_vadv(const Bufs &b, const SphConst &c, int block, hipStream_t s);
// start: WSPH_WARM_* ; writes dp[0] = (d_ii, p1) for a zero start, else (d_ii, p0)
void launch_assemble(const Bufs &b, const SphConst &c, int start_mode, int block, hipStream_t s);
// Feature pass of the learned start (learn.hip): features of every particle,
// optionally the network's start value (net != null) into dp[0].w and aux.z.
struct Weights;
void launch_learned_start(const Bufs &b, const SphConst &c, const Weights *net, hipStream_t s);
// One A p evaluation on dp[src]; writes dp[1-src] = next iterate. slot >= 0
// accumulates the residual of iteration `slot`; slot == -1 marks the
// evaluation of A p0 (counted in DevState::ap_evals).
void launch_j1(const Bufs &b, const SphConst &c, int src, int block, hipStream_t s);
void launch_j2(const Bufs &b, const SphConst &c, int src, int slot, int block, hipStream_t s);
// Stop rule of the device-driven loop (wsph_params' iteration limits and eta).
struct LoopCtl {
  int min_it, max_it, force_it;
  double eta;
};
inline LoopCtl loop_ctl(const wsph_params &P) {
  return LoopCtl{P.min_iters, P.max_iters, P.force_iters, P.eta};
}
// Jacobi iterations k0 .. k0 + iters - 1 with the device-side stop test
// (no-ops once done): one scene tests iteration k - 1 in J1 of iteration k,
// batched scenes run a stop kernel after each J2. ev3 (optional): events
// recorded before J1, between J1 and J2, after J2 of the chunk's first
// iteration.
void launch_jacobi_chunk(const Bufs &b, const SphConst &c, const wsph_params &P, int src0,
                         int k0, int iters, int block, hipStream_t s, hipEvent_t *ev3 = nullptr);
// One scene: stop test of iteration k, the last one the host queued (its
// test has no next J1 to run it); nothing once the loop has ended.
void launch_jacobi_last_test(const Bufs &b, const wsph_params &P, int src0, int k, hipStream_t s);
// Step-end bookkeeping done by the statistics kernel after the kick/drift:
// clock and CFL dt update (IISPH.py:1367-1371) and the cumulative counters.
struct StepEnd {
  int dt_mode;
  double cfl_d, inv_small, dt_max;
  long long *totals;  // cumulative counters (wsph_get_counters)
  int grid;           // persistent grid of the step (0: none)
};
// src < 0: read the final buffer index from DevState
void launch_integrate(const Bufs &b, const SphConst &c, int src, int start_mode, int block,
                      const StepEnd &end, hipStream_t s);
void launch_reset_state(const Bufs &b, hipStream_t s);
// Batched scenes: after the Jacobi loop, move every scene's final iterate
// into dp[final_src] (scenes that stopped early hold it in the other buffer).
void launch_batch_finalize(const Bufs &b, hipStream_t s);
// Persistent cooperative Jacobi loop; returns a hip error code.
hipError_t launch_jacobi_persistent(const Bufs &b, const SphConst &c, const wsph_params &P,
                                    int src, int grid, int block, hipStream_t s);
int persistent_blocks_per_cu(const Bufs &b, const SphConst &c, int block, int *total);
// Field scatter back to original order
void launch_scatter(const void *src, int stride_floats, int ncomp, int src_off, const int *orig,
                    int n, void *dst, hipStream_t s);

// ---- lattice points inside a closed mesh (mesh.hip) ------------------------
struct MeshTri {          // a triangle with its xy projection's data
  double x[3], y[3], z[3];
  double lo[2], hi[2];    // xy bounding box
  int orient;             // +1 counter-clockwise, -1 clockwise, 0 degenerate
  int tiles[4];           // column tiles covered: i-tiles [0]..[1], j-tiles [2]..[3]
};
// number of 16 x 16 column tiles of an nx x ny lattice
long mesh_tiles(int nx, int ny);
// *bad |= 2 unless a[0] < a[1] < ... < a[n - 1]
void launch_mesh_check_axes(const double *a, int n, int *bad, hipStream_t st);
// triangles from vertex/index arrays with their tile ranges, and per-tile
// triangle counts (tile_cnt zeroed by the caller); *bad |= 1 for an index
// outside [0, nv)
void launch_mesh_setup(const double *verts, int nv, const int *tris, int nt, const double *ax,
                       int nx, const double *ay, int ny, MeshTri *out, unsigned *tile_cnt,
                       int *bad, hipStream_t st);
// triangle ids per tile at tile_off (tile_pos zeroed by the caller)
void launch_mesh_bin(const MeshTri *tris, int nt, int ny, const unsigned *tile_off,
                     unsigned *tile_pos, int *ids, hipStream_t st);
// triangles per work item of the column passes
int mesh_chunk();
// out[k] = sum of the counts before k (per > 0: of ceil(count / per)) for n
// counts, in one block; *total = the 64-bit sum
void launch_scan_excl(const unsigned *in, long n, unsigned per, unsigned *out,
                      unsigned long long *total, hipStream_t st);
// work items {tile, first id, count, 0} of the column passes from the per-tile
// triangle counts, their offsets and the tiles' first item
void launch_mesh_items(const unsigned *tile_cnt, const unsigned *tile_off,
                       const unsigned *item_off, long ntiles, int4 *items, hipStream_t st);
// per column (x_i, y_j), c = i * ny + j: crossing counts added to cnt (zbuf
// null) or heights written at off[c] + fill cursor; one block per work item
// {tile, first id, number of ids (<= mesh_chunk())}
void launch_mesh_columns(const MeshTri *tris, const int *ids, const int4 *items, int nitems,
                         const double *ax, int nx, const double *ay, int ny, unsigned *cnt,
                         const unsigned *off, unsigned *fill, double *zbuf, hipStream_t st);
void launch_mesh_classify(const unsigned *off, const unsigned *cnt, double *zbuf, long ncol,
                          const double *az, int nz, unsigned char *inside, hipStream_t st);

// Jacobi stop rule shared by the host loop and the persistent kernel
// (IISPH.py:1200-1207): returns 1 to continue after iteration k.
__host__ __device__ inline int jacobi_continue(int k, unsigned long long sum, unsigned cnt,
                                               int min_it, int max_it, int force_it, double eta,
                                               double *e_out) {
  const double e = cnt == 0 ? 0.0 : ((double)sum / WSPH_RESIDUAL_FIXED_SCALE) / (double)cnt;
  *e_out = e;
  if (force_it > 0) return k + 1 < force_it;
  if (k + 1 >= max_it) return 0;
  if (k + 1 < min_it) return 1;
  return e > eta;
}

}  // namespace wsph
