// capi.hip -- extern "C" entry points of libwsph.so (declared in include/wsph.h).
//
// The solver object owns every device buffer (allocated in wsph_set_*,
// released in wsph_destroy) and one HIP stream at default priority. Particle
// state lives in cell-sorted order; wsph_get_field scatters it back to the
// caller's original order.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "mlp.h"
#include "mlp_dev.h"

using namespace wsph;

static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(WSPH_E_HIP, "%s: %s", #x, hipGetErrorString(e_));     \
  } while (0)

template <typename T>
static hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void **)p, sizeof(T) * count);
}
template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// ---- helper kernels ---------------------------------------------------------
__global__ void k_pack3(const float *__restrict__ x3, int n, float4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = x3 ? make_float4(x3[3 * (long)i], x3[3 * (long)i + 1], x3[3 * (long)i + 2], 0.0f)
              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__global__ void k_iota(int *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
__global__ void k_fill_f32(float *__restrict__ p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Boundary arrays into cell order; orig (null: the input is in original
// order) gives each input slot's original index.
__global__ void k_gather_boundary(const int *__restrict__ perm, int n, const float4 *__restrict__ a,
                                  const float4 *__restrict__ b, const int *__restrict__ orig,
                                  float4 *__restrict__ ao, float4 *__restrict__ bo,
                                  int *__restrict__ orig_o) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int i = perm[f];
  ao[f] = a[i];
  bo[f] = b[i];
  orig_o[f] = orig ? orig[i] : i;
}
// set_rigid_v (IISPH.py:652-660) as launched at IISPH.py:1061-1081: in the
// step that starts at sim_time 0, the moving part takes the rigid velocity.
__global__ void k_set_rigid_v(float4 *__restrict__ bvel, const int *__restrict__ borig, int nb,
                              int first, float vx, float vy, float vz,
                              const DevState *__restrict__ st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb || st->sim_time != 0.0 || borig[k] < first) return;
  bvel[k] = make_float4(vx, vy, vz, bvel[k].w);
}
// update_rigid_x (IISPH.py:663-671): x += v * dt with the step's dt, after
// the fluid drift. w (psi) is kept: the reference never recomputes psi.
__global__ void k_rigid_drift(float4 *__restrict__ bpos, const float4 *__restrict__ bvel,
                              const int *__restrict__ borig, int nb, int first,
                              const DevState *__restrict__ st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb || borig[k] < first) return;
  const float dt = (float)st->dt;
  float4 p = bpos[k];
  const float4 v = bvel[k];
  p.x = p.x + v.x * dt;
  p.y = p.y + v.y * dt;
  p.z = p.z + v.z * dt;
  bpos[k] = p;
}
__global__ void k_copy_boundary(int n, const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                const int *__restrict__ orig, float4 *__restrict__ pos_o,
                                float4 *__restrict__ vel_o, int *__restrict__ orig_o) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  pos_o[k] = pos[k];
  vel_o[k] = vel[k];
  orig_o[k] = orig[k];
}
__global__ void k_gather_by_orig(const float *__restrict__ src_orig, const int *__restrict__ orig,
                                 int n, float *__restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) out[f] = src_orig[orig[f]];
}
__global__ void k_keys4(const float4 *__restrict__ pos, int n, float inv, int d0, int d1, int d2,
                        int *__restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int dims[3] = {d0, d1, d2};
  const float4 p = pos[i];
  keys[i] = linear_key(p.x, p.y, p.z, inv, dims);
}
__global__ void k_export_lists(const int *__restrict__ nbr, const int *__restrict__ cnt, int n,
                               const int *__restrict__ orig, const int *__restrict__ other_orig,
                               int cap_in, int cap_out, int *__restrict__ cnt_o,
                               int *__restrict__ idx_o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long o = orig[i];
  const int c = cnt[i];
  cnt_o[o] = c;
  const int *row = nbr + list_base(i, cap_in);
  for (int k = 0; k < c && k < cap_out; ++k) idx_o[o * cap_out + k] = other_orig[row[(long)k * kListLanes]];
}
__global__ void k_solve_result(DevState *st, int iters, double residual, int converged,
                               int ap_evals) {
  st->iterations = iters;
  st->residual = residual;
  st->converged = converged;
  st->ap_evals += ap_evals;
}
__global__ void k_set_dt(DevState *st, double dt, double t) {
  st->dt = dt;
  st->sim_time = t;
}

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- solver object ------------------------------------------------------------
struct GridBufs {
  long ncell = 0;
  unsigned *count = nullptr, *start = nullptr;
  std::vector<unsigned *> scan;
  int dims[3] = {0, 0, 0};
};

// ---- pressure predictor network ---------------------------------------------
struct wsph_mlp {
  MlpDev dev;
  float *adam_m = nullptr, *adam_v = nullptr, *loss = nullptr, *grad = nullptr;
  hipStream_t stream = nullptr;
  int step = 0;
};

struct wsph_solver {
  wsph_params P;
  SphConst C;
  hipStream_t stream = nullptr;
  int n = 0, nb = 0, cap_n = 0;
  float *wt = nullptr, *gt = nullptr;
  bool lds_tables = false;
  int lds_per_cu = 0;
  DevState *st = nullptr;
  DevState *st_host = nullptr;   // pinned mirror
  long long *totals = nullptr;   // device counters
  // fluid (double-buffered across the per-step reorder)
  float4 *pos[2] = {nullptr, nullptr}, *vel[2] = {nullptr, nullptr};
  float *pprev[2] = {nullptr, nullptr};
  int *orig[2] = {nullptr, nullptr};
  int cur = 0;
  float4 *vadv = nullptr, *dp[2] = {nullptr, nullptr}, *sap = nullptr, *aux = nullptr;
  float *ap = nullptr, *pstart = nullptr, *p_ext = nullptr, *pfinal = nullptr;
  unsigned *vmax_part = nullptr, *ymax_part = nullptr;
  double *ke_part = nullptr, *y_part = nullptr;
  StepRecord *history = nullptr;
  unsigned long long *res_part = nullptr;
  int *list_max = nullptr;  // kListMaxSlots slots of the list pass's largest counts
  unsigned *cnt_part = nullptr;
  bool have_p_ext = false;
  int *slot = nullptr, *rank = nullptr, *tmp = nullptr, *perm = nullptr;
  int *nbr_ff = nullptr, *cnt_ff = nullptr, *nbr_fs = nullptr, *cnt_fs = nullptr;
  GridBufs fg, bg;
  // boundary
  float4 *bpos = nullptr, *bvel = nullptr;
  int *borig = nullptr;
  unsigned *bnear = nullptr;  // boundary slots near a boundary particle (static boundary)
  // rigid motion of the boundary particles with original index >= rigid_first
  // (-1: static boundary), and the buffers of the per-step boundary re-sort
  int rigid_first = -1;
  float rigid_v[3] = {0.0f, 0.0f, 0.0f};
  // the velocity set has not been launched yet: only the first step after
  // wsph_set_rigid launches it (sim_time grows from then on), eagerly
  bool rigid_v_pending = false;
  float4 *rb_pos = nullptr, *rb_vel = nullptr;
  int *rb_orig = nullptr, *rb_slot = nullptr, *rb_rank = nullptr, *rb_tmp = nullptr,
      *rb_perm = nullptr;
  int final_dp = 0;
  bool searched = false, advected = false, solved = false;
  // timing: a ring of per-step event sets, read back lazily so that timed
  // steps need no host synchronisation
  int timing = 0;  // 0 off, 1 phases and kernels, 2 kernels of every kKernelEvery-th step
  // per step: 5 phase boundaries, then 3 events around J1 and J2 of the
  // first Jacobi iteration (device driver; that iteration always does work)
  static constexpr int kRing = 64, kEv = 8;
  static constexpr int kKernelEvery = 4;  // timing mode 2: J1/J2 events on every 4th step
  hipEvent_t ev[kRing * kEv] = {};
  bool kev_rec[kRing] = {};
  double k_sum[2] = {0, 0};
  long k_steps = 0;
  double k_pair = 0;  // J1 + J2 between two events (the first before J1, the second after J2)
  long k_pairs = 0;
  long t_recorded = 0, t_processed = 0;
  double t_sum[kEv] = {0};
  int persistent_grid = 0;
  // device-driven Jacobi loop: pinned host copy of DevState::done_tag, stored
  // by the device's stop test
  int *done_host = nullptr;
  unsigned long long *prog_host = nullptr;
  int step_count = 0;
  // captured step graphs, one per fluid buffer parity at step start
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int graph_src = 0;  // Jacobi source buffer after the start value (graph mode)
  // batched scenes (nscenes == 0: one scene described by the params)
  int nscenes = 0;
  std::vector<int> scene_nf, scene_nb;
  std::vector<SceneGrid> scene_grid;
  SceneGrid *d_scenes = nullptr;
  int *sid[2] = {nullptr, nullptr};
  int *bsid = nullptr;
  unsigned long long *sc_sum = nullptr;
  unsigned *sc_cnt = nullptr;
  int *sc_done = nullptr, *sc_iters = nullptr, *sc_final = nullptr;
  double *sc_res = nullptr;
  // learned start and training-pair logging
  wsph_mlp *net = nullptr;
  bool log_pairs = false, have_pairs = false;
  float4 *seed = nullptr;
  float *feat = nullptr, *target = nullptr;
  double *err_part = nullptr;
};

// Smallest float32 d2 >= 0 with sqrtf(d2) >= h: then sqrtf(d2) < h exactly
// when d2 < threshold, because rounded sqrt is monotone.
static float radius_sq_threshold(float h) {
  uint32_t lo = 0, hi = 0x7f7fffffu;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    float s;
    memcpy(&s, &mid, sizeof(s));
    if (sqrtf(s) >= h)
      hi = mid;
    else
      lo = mid + 1;
  }
  float r;
  memcpy(&r, &lo, sizeof(r));
  return r;
}

static long grid_cells(const int *dims, int layout) {
  if (layout == WSPH_LAYOUT_MORTON) return (long)morton3(dims[0] - 1, dims[1] - 1, dims[2] - 1) + 1;
  return (long)dims[0] * dims[1] * dims[2];
}

static int alloc_grid(GridBufs &g, const int *dims, int layout, long ncell = -1) {
  for (auto p : g.scan) dfree(p);
  g.scan.clear();
  dfree(g.count);
  dfree(g.start);
  memcpy(g.dims, dims, sizeof(g.dims));
  g.ncell = ncell >= 0 ? ncell : grid_cells(dims, layout);
  HIP_TRY(dalloc(&g.count, g.ncell));
  HIP_TRY(dalloc(&g.start, g.ncell));
  long sizes[8];
  const int L = scan_levels_needed(g.ncell, sizes, 8);
  for (int l = 0; l < L; ++l)
    for (int r = 0; r < 2; ++r) {
      unsigned *p = nullptr;
      HIP_TRY(dalloc(&p, sizes[l]));
      g.scan.push_back(p);
    }
  g.scan.push_back(nullptr);
  g.scan.push_back(nullptr);
  return WSPH_OK;
}

static Bufs make_bufs(wsph_solver *s) {
  Bufs b;
  memset(&b, 0, sizeof(b));
  b.n = s->n;
  b.nb = s->nb;
  b.pos = s->pos[s->cur];
  b.vel = s->vel[s->cur];
  b.vadv = s->vadv;
  b.dp[0] = s->dp[0];
  b.dp[1] = s->dp[1];
  b.sap = s->sap;
  b.aux = s->aux;
  b.ap = s->ap;
  b.pprev = s->pprev[s->cur];
  b.pstart = s->pstart;
  b.pfinal = s->pfinal;
  b.vmax_part = s->vmax_part;
  b.ymax_part = s->ymax_part;
  b.ke_part = s->ke_part;
  b.y_part = s->y_part;
  b.history = s->history;
  b.res_part = s->res_part;
  b.list_max = s->list_max;
  b.cnt_part = s->cnt_part;
  b.fstart = s->fg.start;
  b.fcount = s->fg.count;
  b.nbr_ff = s->nbr_ff;
  b.cnt_ff = s->cnt_ff;
  b.nbr_fs = s->nbr_fs;
  b.cnt_fs = s->cnt_fs;
  b.cap_ff = s->P.max_ff;
  b.cap_fs = s->P.max_fs;
  b.bpos = s->bpos;
  b.bvel = s->bvel;
  b.bstart = s->bg.start;
  b.bcount = s->bg.count;
  // a moving boundary is re-sorted every step; its queries search every cell
  b.bnear = s->rigid_first < 0 ? s->bnear : nullptr;
  b.wt = s->wt;
  b.gt = s->gt;
  b.lds_tables = s->lds_tables;
  b.lds_per_cu = s->lds_per_cu;
  b.st = s->st;
  const bool learned = s->P.warm_start == WSPH_WARM_LEARNED;
  if (learned || s->log_pairs) b.seed = s->seed;
  if (s->log_pairs) {
    b.feat = s->feat;
    b.target = s->target;
  }
  b.err_part = s->err_part;
  b.done_host = s->done_host;
  b.prog_host = s->prog_host;
  b.arbiter = learned && s->net && s->err_part;
  if (s->nscenes > 0) {
    b.scenes = s->d_scenes;
    b.sid = s->sid[s->cur];
    b.bsid = s->bsid;
    b.nscenes = s->nscenes;
    b.sc_sum = s->sc_sum;
    b.sc_cnt = s->sc_cnt;
    b.sc_done = s->sc_done;
    b.sc_iters = s->sc_iters;
    b.sc_final = s->sc_final;
    b.sc_res = s->sc_res;
  }
  return b;
}

// Scale of a_ii for the predictor's features: a particle inside a resting
// lattice of spacing d has d_ii = 0 (symmetric neighbourhood) and
// a_ii = -dt^2 m^2 / rho0^2 * S with S = sum_j |grad W(x_j)|^2 over the lattice
// points within h (cubic spline of sph_funcs.py:56-82). Returns
// -rho0^2 / (m^2 S), so that a_ii / a_ref = a_ii * k1 / dt^2.
static float feature_k1(const wsph_params &P) {
  const double h = P.h, d = P.spacing, sigma = 8.0 / (M_PI * h * h * h);
  const int R = (int)(h / d) + 1;
  double S = 0.0;
  for (int a = -R; a <= R; ++a)
    for (int b = -R; b <= R; ++b)
      for (int c = -R; c <= R; ++c) {
        const double r = d * sqrt((double)(a * a + b * b + c * c));
        if (r <= 0.0 || r >= h) continue;
        const double q = r / h;
        const double g = q <= 0.5 ? (12.0 * q - 18.0 * q * q) * sigma / h
                                  : 6.0 * (1.0 - q) * (1.0 - q) * sigma / h;
        S += g * g;
      }
  if (S <= 0.0) return 0.0f;
  return (float)(-(P.rho0 * P.rho0) / (P.mass * P.mass * S));
}

static int check_dims(const int *d, const char *what) {
  for (int k = 0; k < 3; ++k)
    if (d[k] < 1 || d[k] > 1024) return fail(WSPH_E_ARG, "%s dims must be in [1, 1024]", what);
  return WSPH_OK;
}

// ---- wsph_mesh_inside_noalloc: the same classification with no allocation --
// Every scratch buffer of the classification lives in one file-scope device
// array (each GPU the library runs on holds its own copy, at its own address,
// until the process ends). Host code never uses the array's name as a device
// address: hipGetSymbolAddress gives the address on each device once, and
// the per-device pointer is kept.
constexpr size_t kMeshScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kMeshScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[kScratchDevices] = {};

static int scratch_base(unsigned char **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return fail(WSPH_E_UNSUPPORTED, "device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(WSPH_E_HIP, "no device address for the mesh scratch array");
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_dev[dev];
  return WSPH_OK;
}

// Consecutive 256-byte aligned regions of the scratch array; with base null
// it only adds up the sizes.
struct ScratchCarve {
  unsigned char *base;
  size_t used = 0;
  bool over = false;
  template <typename T>
  T *take(size_t count) {
    const size_t off = (used + 255) & ~size_t(255);
    const size_t bytes = sizeof(T) * (count > 0 ? count : 1);
    if (off > kMeshScratchBytes || bytes > kMeshScratchBytes - off) {
      over = true;
      return nullptr;
    }
    used = off + bytes;
    return base ? reinterpret_cast<T *>(base + off) : nullptr;
  }
};

extern "C" {

const char *wsph_last_error(void) { return g_err; }
float wsph_radius_sq_threshold(float h) { return radius_sq_threshold(h); }
int wsph_abi_version(void) { return WSPH_ABI_VERSION; }

int wsph_blocks_per_cu(int block_threads, int vgprs, int sgprs, int lds_bytes, int cu_lds_bytes,
                       int waves_per_simd) {
  if (block_threads < 1 || block_threads > 1024 || vgprs < 0 || sgprs > 112 || lds_bytes < 0 ||
      cu_lds_bytes <= 0 || waves_per_simd <= 0)
    return fail(WSPH_E_ARG, "bad resource figures");
  constexpr int kSimds = 4, kVgprsPerLane = 512, kVgprGranule = 8, kLdsGranule = 512,
                kMaxGroups = 16, kSgprsPerSimd = 800, kSgprGranule = 16, kSgprExtra = 16,
                kSgprMax = 112;
  const int waves = (block_threads + 63) / 64;
  // a block's waves spread over the 4 SIMDs: at most ceil(waves / 4) on each
  const int per_simd = (waves + kSimds - 1) / kSimds;
  int wps = waves_per_simd;
  const int v = (vgprs + kVgprGranule - 1) / kVgprGranule * kVgprGranule;
  if (v > 0) wps = wps < kVgprsPerLane / v ? wps : kVgprsPerLane / v;
  const int sg = sgprs > 0 ? sgprs : kSgprMax;
  const int s = (sg + kSgprGranule - 1) / kSgprGranule * kSgprGranule + kSgprExtra;
  wps = wps < kSgprsPerSimd / s ? wps : kSgprsPerSimd / s;
  int blocks = wps / per_simd;
  const int l = (lds_bytes + kLdsGranule - 1) / kLdsGranule * kLdsGranule;
  if (l > 0) blocks = blocks < cu_lds_bytes / l ? blocks : cu_lds_bytes / l;
  return blocks < kMaxGroups ? blocks : kMaxGroups;
}

int wsph_selftest_arith(unsigned long long *out) {
  if (!out) return fail(WSPH_E_ARG, "null argument");
  unsigned long long *d = nullptr;
  HIP_TRY(hipMalloc(&d, 4 * sizeof(unsigned long long)));
  hipError_t e = hipMemset(d, 0, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) {
    launch_selftest_arith(d, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  HIP_TRY(e);
  return WSPH_OK;
}

int wsph_create(const wsph_params *params, wsph_solver **out) {
  if (!params || !out) return fail(WSPH_E_ARG, "null argument");
  const wsph_params &P = *params;
  if (!(P.h > 0 && P.spacing > 0 && P.rho0 > 0 && P.mass > 0 && P.dt > 0))
    return fail(WSPH_E_ARG, "h, spacing, rho0, mass and dt must be positive");
  if (P.max_iters < 1 || P.max_iters > kMaxIters || P.force_iters < 0 ||
      P.force_iters > kMaxIters || P.min_iters < 0)
    return fail(WSPH_E_ARG, "iteration bounds must satisfy 1 <= max_iters <= %d", kMaxIters);
  if (P.max_ff < 4 || P.max_fs < 4 || (P.max_ff & 3) || (P.max_fs & 3))
    return fail(WSPH_E_ARG, "neighbour capacities must be positive multiples of 4");
  if (P.block < 64 || P.block > 1024 || (P.block & 63))
    return fail(WSPH_E_ARG, "block must be a multiple of 64 in [64, 1024]");
  if (P.cell_layout != WSPH_LAYOUT_LINEAR && P.cell_layout != WSPH_LAYOUT_MORTON)
    return fail(WSPH_E_ARG, "unknown cell layout");
  if (P.warm_start < WSPH_WARM_ZERO || P.warm_start > WSPH_WARM_LEARNED)
    return fail(WSPH_E_ARG, "unknown warm start mode");
  int rc;
  if ((rc = check_dims(P.grid_dims, "grid")) || (rc = check_dims(P.bgrid_dims, "boundary grid")))
    return rc;
  wsph_solver *s = new wsph_solver();
  s->P = P;
  SphConst &C = s->C;
  memset(&C, 0, sizeof(C));
  C.h = (float)P.h;
  C.r2_lt = radius_sq_threshold(C.h);
  C.cell_inv = 1.0f / C.h;
  C.mass = (float)P.mass;
  C.rho0 = (float)P.rho0;
  C.omega = (float)P.omega;
  C.inv_small = (float)P.inv_small;
  C.g_y = -(float)P.gravity;
  C.vis_mu = (float)P.vis_mu;
  C.xsph = (float)P.xsph;
  memcpy(C.fdims, P.grid_dims, sizeof(C.fdims));
  memcpy(C.bdims, P.bgrid_dims, sizeof(C.bdims));
  C.layout = P.cell_layout;
  C.feat_k1 = feature_k1(P);
  const double g_abs = fabs(P.gravity) > 1.0 ? fabs(P.gravity) : 1.0;
  C.pref = (float)(P.rho0 * g_abs * 0.1);
  C.inv_pref = 1.0f / C.pref;
  *out = nullptr;
#define CREATE_TRY(x)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      int r_ = fail(WSPH_E_HIP, "%s: %s", #x, hipGetErrorString(e_));                        \
      wsph_destroy(s);                                                                       \
      return r_;                                                                             \
    }                                                                                        \
  } while (0)
  int dev = 0;
  CREATE_TRY(hipGetDevice(&dev));
  CREATE_TRY(hipDeviceGetAttribute(&s->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  CREATE_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  CREATE_TRY(dalloc(&s->st, 1));
  CREATE_TRY(dalloc(&s->list_max, (size_t)kListMaxSlots * kListMaxStride));
  CREATE_TRY(hipMemsetAsync(s->list_max, 0, sizeof(int) * kListMaxSlots * kListMaxStride, s->stream));
  CREATE_TRY(hipHostMalloc((void **)&s->st_host, sizeof(DevState), hipHostMallocDefault));
  CREATE_TRY(dalloc(&s->totals, 8));
  CREATE_TRY(dalloc(&s->history, kHistory));
  CREATE_TRY(hipMemsetAsync(s->totals, 0, sizeof(long long) * 8, s->stream));
  CREATE_TRY(hipMemsetAsync(s->st, 0, sizeof(DevState), s->stream));
  for (auto &e : s->ev) CREATE_TRY(hipEventCreate(&e));
  // coherent (fine-grained) pinned words: the device's stores reach the host
  // while its kernels run, without a release at a kernel or event boundary
  CREATE_TRY(hipHostMalloc((void **)&s->done_host, sizeof(int), hipHostMallocCoherent));
  *s->done_host = -1;
  CREATE_TRY(hipHostMalloc((void **)&s->prog_host, sizeof(unsigned long long), hipHostMallocCoherent));
  *s->prog_host = ~0ull;
  k_set_dt<<<1, 1, 0, s->stream>>>(s->st, P.dt, 0.0);
  CREATE_TRY(hipGetLastError());
  int rcg = alloc_grid(s->fg, P.grid_dims, P.cell_layout);
  if (!rcg) rcg = alloc_grid(s->bg, P.bgrid_dims, P.cell_layout);
  if (rcg) {
    wsph_destroy(s);
    return rcg;
  }
  CREATE_TRY(hipStreamSynchronize(s->stream));
#undef CREATE_TRY
  *out = s;
  return WSPH_OK;
}

static void drop_graphs(wsph_solver *s) {
  for (int k = 0; k < 2; ++k) {
    if (s->gexec[k]) (void)hipGraphExecDestroy(s->gexec[k]);
    if (s->graph[k]) (void)hipGraphDestroy(s->graph[k]);
    s->gexec[k] = nullptr;
    s->graph[k] = nullptr;
  }
}

static void free_rigid(wsph_solver *s) {
  drop_graphs(s);
  // the boundary may have moved: its near-boundary marks no longer hold
  if (s->rigid_first >= 0) dfree(s->bnear);
  s->rigid_first = -1;
  s->rigid_v_pending = false;
  dfree(s->rb_pos);
  dfree(s->rb_vel);
  dfree(s->rb_orig);
  dfree(s->rb_slot);
  dfree(s->rb_rank);
  dfree(s->rb_tmp);
  dfree(s->rb_perm);
}

static void free_fluid(wsph_solver *s) {
  drop_graphs(s);
  for (int k = 0; k < 2; ++k) {
    dfree(s->pos[k]);
    dfree(s->vel[k]);
    dfree(s->pprev[k]);
    dfree(s->orig[k]);
    dfree(s->dp[k]);
  }
  dfree(s->vadv);
  dfree(s->sap);
  dfree(s->aux);
  dfree(s->ap);
  dfree(s->pstart);
  dfree(s->p_ext);
  dfree(s->pfinal);
  dfree(s->vmax_part);
  dfree(s->ymax_part);
  dfree(s->ke_part);
  dfree(s->y_part);
  dfree(s->res_part);
  dfree(s->cnt_part);
  dfree(s->slot);
  dfree(s->rank);
  dfree(s->tmp);
  dfree(s->perm);
  dfree(s->nbr_ff);
  dfree(s->cnt_ff);
  dfree(s->nbr_fs);
  dfree(s->cnt_fs);
  dfree(s->sid[0]);
  dfree(s->sid[1]);
  dfree(s->seed);
  dfree(s->feat);
  dfree(s->target);
  dfree(s->err_part);
  s->have_pairs = false;
  s->n = 0;
}

// Buffers of the learned start / pair logging, allocated on first use.
static int ensure_learn_bufs(wsph_solver *s) {
  const size_t n = (size_t)s->cap_n;
  if (!s->seed) HIP_TRY(dalloc(&s->seed, n));
  if (!s->err_part) HIP_TRY(dalloc(&s->err_part, 2 * ((n + 63) / 64 + 1)));
  if (s->log_pairs) {
    if (!s->feat) HIP_TRY(dalloc(&s->feat, n * WSPH_MLP_IN));
    if (!s->target) HIP_TRY(dalloc(&s->target, n));
  }
  return WSPH_OK;
}

int wsph_destroy(wsph_solver *s) {
  if (!s) return WSPH_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_fluid(s);
  free_rigid(s);
  dfree(s->bpos);
  dfree(s->bvel);
  dfree(s->borig);
  dfree(s->bnear);
  dfree(s->bsid);
  dfree(s->d_scenes);
  dfree(s->sc_sum);
  dfree(s->sc_cnt);
  dfree(s->sc_done);
  dfree(s->sc_iters);
  dfree(s->sc_final);
  dfree(s->sc_res);
  dfree(s->wt);
  dfree(s->gt);
  dfree(s->st);
  dfree(s->list_max);
  dfree(s->totals);
  dfree(s->history);
  if (s->st_host) (void)hipHostFree(s->st_host);
  for (GridBufs *g : {&s->fg, &s->bg}) {
    dfree(g->count);
    dfree(g->start);
    for (auto p : g->scan) dfree(p);
  }
  for (auto &e : s->ev)
    if (e) (void)hipEventDestroy(e);
  if (s->done_host) (void)hipHostFree(s->done_host);
  if (s->prog_host) (void)hipHostFree(s->prog_host);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return WSPH_OK;
}

int wsph_set_tables(wsph_solver *s, const float *w_table, int w_size, const float *grad_table,
                    int grad_size) {
  if (!s || !w_table || !grad_table || w_size < 2 || grad_size < 2)
    return fail(WSPH_E_ARG, "tables need at least 2 samples");
  drop_graphs(s);
  dfree(s->wt);
  dfree(s->gt);
  HIP_TRY(dalloc(&s->wt, w_size));
  HIP_TRY(dalloc(&s->gt, grad_size));
  HIP_TRY(hipMemcpyAsync(s->wt, w_table, sizeof(float) * w_size, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->gt, grad_table, sizeof(float) * grad_size, hipMemcpyHostToDevice,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->C.nw = w_size;
  s->C.ng = grad_size;
  const float h_inv = (float)(1.0 / s->P.h);
  s->C.w_scale = (float)(w_size - 1) * h_inv;
  s->C.g_scale = (float)(grad_size - 1) * h_inv;
  s->lds_tables = (w_size + grad_size) <= kMaxLdsTable;
  return WSPH_OK;
}

// Sort points (float4, current order = input order) into cell order.
static int sort_points(wsph_solver *s, GridBufs &g, const float4 *pos, const int *orig, int n,
                       int *slot, int *rank, int *tmp, int *perm, const int *sid = nullptr,
                       int boundary = 0) {
  GridArgs a;
  a.sid = s->nscenes > 0 ? sid : nullptr;
  a.scenes = s->nscenes > 0 ? s->d_scenes : nullptr;
  a.boundary = boundary;
  a.n = n;
  a.ncell = g.ncell;
  a.inv = s->C.cell_inv;
  memcpy(a.dims, g.dims, sizeof(a.dims));
  a.layout = s->P.cell_layout;
  a.pos = pos;
  a.orig = orig;
  a.slot = slot;
  a.rank = rank;
  a.tmp = tmp;
  a.perm = perm;
  a.count = g.count;
  a.start = g.start;
  a.scan_bufs = g.scan.data();
  sort_into_cells(a, s->stream);
  HIP_TRY(hipGetLastError());
  return WSPH_OK;
}

int wsph_set_boundary(wsph_solver *s, const float *xb, const float *vb, int nb) {
  if (!s || nb < 0 || (nb > 0 && !xb)) return fail(WSPH_E_ARG, "bad boundary arguments");
  if (nb > kMaxParticles) return fail(WSPH_E_CAPACITY, "%d boundary particles (at most %d)", nb, kMaxParticles);
  if (!s->wt) return fail(WSPH_E_ARG, "call wsph_set_tables first");
  drop_graphs(s);
  free_rigid(s);
  dfree(s->bpos);
  dfree(s->bvel);
  dfree(s->borig);
  dfree(s->bnear);
  s->nb = nb;
  HIP_TRY(dalloc(&s->bpos, nb + kPointPad));
  HIP_TRY(dalloc(&s->bvel, nb));
  HIP_TRY(dalloc(&s->borig, nb));
  HIP_TRY(dalloc(&s->bnear, s->bg.ncell));
  float4 *p4 = nullptr, *v4 = nullptr;
  int *id = nullptr, *sl = nullptr, *rk = nullptr, *tp = nullptr, *pm = nullptr, *bs = nullptr;
  int rc = WSPH_OK;
  if (s->nscenes > 0) {
    long total = 0;
    for (int v : s->scene_nb) total += v;
    if (total != nb) return fail(WSPH_E_ARG, "boundary count %d differs from the scenes' %ld", nb, total);
    dfree(s->bsid);
    HIP_TRY(dalloc(&s->bsid, nb));
    if (nb > 0) {
      std::vector<int> h(nb);
      int o = 0;
      for (int k = 0; k < s->nscenes; ++k)
        for (int q = 0; q < s->scene_nb[k]; ++q) h[o++] = k;
      HIP_TRY(dalloc(&bs, nb));
      HIP_TRY(hipMemcpy(bs, h.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
    }
  }
  if (nb > 0) {
    HIP_TRY(dalloc(&p4, nb));
    HIP_TRY(dalloc(&v4, nb));
    HIP_TRY(dalloc(&id, nb));
    HIP_TRY(dalloc(&sl, nb));
    HIP_TRY(dalloc(&rk, nb));
    HIP_TRY(dalloc(&tp, nb));
    HIP_TRY(dalloc(&pm, nb));
    k_pack3<<<nblk(nb, 256), 256, 0, s->stream>>>(xb, nb, p4);
    k_pack3<<<nblk(nb, 256), 256, 0, s->stream>>>(vb, nb, v4);
    k_iota<<<nblk(nb, 256), 256, 0, s->stream>>>(id, nb);
  }
  rc = sort_points(s, s->bg, p4, id, nb, sl, rk, tp, pm, bs, 1);
  if (!rc && nb > 0) {
    k_gather_boundary<<<nblk(nb, 256), 256, 0, s->stream>>>(pm, nb, p4, v4, nullptr, s->bpos,
                                                            s->bvel, s->borig);
    if (bs) gather_i32(pm, nb, bs, s->bsid, s->stream);
    Bufs b = make_bufs(s);
    launch_boundary_psi(b, s->C, s->stream);
    fill_u32(s->bnear, 0u, s->bg.ncell, s->stream);
    launch_mark_bnear(b, s->C, s->bnear, s->stream);
    if (hipGetLastError() != hipSuccess) rc = fail(WSPH_E_HIP, "boundary psi launch failed");
  }
  hipError_t e = hipStreamSynchronize(s->stream);
  dfree(p4);
  dfree(v4);
  dfree(id);
  dfree(sl);
  dfree(rk);
  dfree(tp);
  dfree(pm);
  dfree(bs);
  if (e != hipSuccess) return fail(WSPH_E_HIP, "boundary setup: %s", hipGetErrorString(e));
  return rc;
}

int wsph_set_fluid(wsph_solver *s, const float *x, const float *v, int n) {
  if (!s || n < 0 || (n > 0 && !x)) return fail(WSPH_E_ARG, "bad fluid arguments");
  if (n > kMaxParticles) return fail(WSPH_E_CAPACITY, "%d fluid particles (at most %d)", n, kMaxParticles);
  if (!s->wt) return fail(WSPH_E_ARG, "call wsph_set_tables first");
  if (!s->bpos) return fail(WSPH_E_ARG, "call wsph_set_boundary first");
  if (s->nscenes > 0) {
    long total = 0;
    for (int v : s->scene_nf) total += v;
    if (total != n) return fail(WSPH_E_ARG, "fluid count %d differs from the scenes' %ld", n, total);
  }
  if (n != s->cap_n || !s->pos[0]) {  // (a first call with n = 0 allocates too)
    free_fluid(s);
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(dalloc(&s->pos[k], n + kPointPad));
      HIP_TRY(dalloc(&s->vel[k], n));
      HIP_TRY(dalloc(&s->pprev[k], n));
      HIP_TRY(dalloc(&s->orig[k], n));
      HIP_TRY(dalloc(&s->dp[k], n));
    }
    HIP_TRY(dalloc(&s->vadv, n));
    HIP_TRY(dalloc(&s->sap, n));
    HIP_TRY(dalloc(&s->aux, n));
    HIP_TRY(dalloc(&s->ap, n));
    HIP_TRY(dalloc(&s->pstart, n));
    HIP_TRY(dalloc(&s->pfinal, n));
    HIP_TRY(dalloc(&s->vmax_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->ymax_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->ke_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->y_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->res_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->cnt_part, (size_t)(n + 63) / 64 + 1));
    HIP_TRY(dalloc(&s->p_ext, n));
    HIP_TRY(dalloc(&s->slot, n));
    HIP_TRY(dalloc(&s->rank, n));
    HIP_TRY(dalloc(&s->tmp, n));
    HIP_TRY(dalloc(&s->perm, n));
    HIP_TRY(dalloc(&s->cnt_ff, n));
    HIP_TRY(dalloc(&s->cnt_fs, n));
    HIP_TRY(dalloc(&s->nbr_ff, list_entries(n, s->P.max_ff)));
    HIP_TRY(dalloc(&s->nbr_fs, list_entries(n, s->P.max_fs)));
    s->cap_n = n;
  }
  if (s->nscenes > 0 && n > 0) {
    dfree(s->sid[0]);
    dfree(s->sid[1]);
    HIP_TRY(dalloc(&s->sid[0], n));
    HIP_TRY(dalloc(&s->sid[1], n));
    std::vector<int> h(n);
    int o = 0;
    for (int k = 0; k < s->nscenes; ++k)
      for (int q = 0; q < s->scene_nf[k]; ++q) h[o++] = k;
    HIP_TRY(hipMemcpy(s->sid[0], h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  }
  s->n = n;
  s->cur = 0;
  s->have_p_ext = false;
  if (n > 0) {
    k_pack3<<<nblk(n, 256), 256, 0, s->stream>>>(x, n, s->pos[0]);
    k_pack3<<<nblk(n, 256), 256, 0, s->stream>>>(v, n, s->vel[0]);
    k_iota<<<nblk(n, 256), 256, 0, s->stream>>>(s->orig[0], n);
    k_fill_f32<<<nblk(n, 256), 256, 0, s->stream>>>(s->pprev[0], n, 0.0f);
    k_fill_f32<<<nblk(n, 256), 256, 0, s->stream>>>(s->p_ext, n, 0.0f);
    k_fill_f32<<<nblk(n, 256), 256, 0, s->stream>>>(s->pfinal, n, 0.0f);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->searched = s->advected = s->solved = false;
  return WSPH_OK;
}

int wsph_set_pressure(wsph_solver *s, const float *p) {
  if (!s || !p) return fail(WSPH_E_ARG, "null argument");
  if (s->n > 0)
    HIP_TRY(hipMemcpyAsync(s->p_ext, p, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream));
  s->have_p_ext = true;
  return WSPH_OK;
}

static void rec(wsph_solver *s, int k) {
  if (s->timing == 1)
    (void)hipEventRecord(s->ev[(s->t_recorded % wsph_solver::kRing) * wsph_solver::kEv + k],
                         s->stream);
}

// Fold the oldest recorded step into the phase sums.
static void timing_fold_one(wsph_solver *s) {
  const int base = (int)(s->t_processed % wsph_solver::kRing) * wsph_solver::kEv;
  const int slot = (int)(s->t_processed % wsph_solver::kRing);
  float ms;
  if (s->timing == 1) {
    (void)hipEventSynchronize(s->ev[base + 4]);
    for (int k = 0; k < 4; ++k) {
      if (hipEventElapsedTime(&ms, s->ev[base + k], s->ev[base + k + 1]) == hipSuccess)
        s->t_sum[k] += ms;
    }
    if (hipEventElapsedTime(&ms, s->ev[base], s->ev[base + 4]) == hipSuccess) s->t_sum[4] += ms;
  } else if (s->kev_rec[slot]) {
    (void)hipEventSynchronize(s->ev[base + 7]);
  }
  if (s->kev_rec[slot]) {
    float m1, m2;
    if (s->timing == 1 &&
        hipEventElapsedTime(&m1, s->ev[base + 5], s->ev[base + 6]) == hipSuccess &&
        hipEventElapsedTime(&m2, s->ev[base + 6], s->ev[base + 7]) == hipSuccess) {
      s->k_sum[0] += m1;
      s->k_sum[1] += m2;
      s->k_steps++;
    }
    if (hipEventElapsedTime(&m1, s->ev[base + 5], s->ev[base + 7]) == hipSuccess) {
      s->k_pair += m1;
      s->k_pairs++;
    }
    s->kev_rec[slot] = false;
  }
  s->t_processed++;
}

// Events around the first Jacobi iteration's J1 and J2 of the step being
// recorded (null when timing is off).
static void kernel_events(wsph_solver *s, hipEvent_t ev3[3]) {
  if (!s->timing || (s->timing == 2 && s->t_recorded % wsph_solver::kKernelEvery != 0)) {
    ev3[0] = ev3[1] = ev3[2] = nullptr;
    return;
  }
  const int slot = (int)(s->t_recorded % wsph_solver::kRing);
  const int base = slot * wsph_solver::kEv;
  for (int k = 0; k < 3; ++k) ev3[k] = s->ev[base + 5 + k];
  if (s->timing == 2) ev3[1] = nullptr;  // the pair only: one event less inside it
  s->kev_rec[slot] = true;
}

// DYNAMIC_SCENE's boundary grid rebuild (IISPH.py:1415-1416): the moved
// boundary, in its current cell order, is sorted again into cells, each
// cell ordered by original index as at set-up.
static int resort_boundary(wsph_solver *s) {
  const int nb = s->nb;
  if (s->rigid_first < 0 || nb == 0) return WSPH_OK;
  k_copy_boundary<<<nblk(nb, 256), 256, 0, s->stream>>>(nb, s->bpos, s->bvel, s->borig, s->rb_pos,
                                                         s->rb_vel, s->rb_orig);
  int rc = sort_points(s, s->bg, s->rb_pos, s->rb_orig, nb, s->rb_slot, s->rb_rank, s->rb_tmp,
                       s->rb_perm, nullptr, 1);
  if (rc) return rc;
  k_gather_boundary<<<nblk(nb, 256), 256, 0, s->stream>>>(s->rb_perm, nb, s->rb_pos, s->rb_vel,
                                                          s->rb_orig, s->bpos, s->bvel, s->borig);
  HIP_TRY(hipGetLastError());
  return WSPH_OK;
}

static void launch_rigid_v(wsph_solver *s) {
  if (s->rigid_first < 0 || s->nb == 0 || !s->rigid_v_pending) return;
  s->rigid_v_pending = false;
  k_set_rigid_v<<<nblk(s->nb, 256), 256, 0, s->stream>>>(s->bvel, s->borig, s->nb, s->rigid_first,
                                                         s->rigid_v[0], s->rigid_v[1],
                                                         s->rigid_v[2], s->st);
}

static void launch_rigid_drift(wsph_solver *s) {
  if (s->rigid_first < 0 || s->nb == 0) return;
  k_rigid_drift<<<nblk(s->nb, 256), 256, 0, s->stream>>>(s->bpos, s->bvel, s->borig, s->nb,
                                                         s->rigid_first, s->st);
}

int wsph_set_rigid(wsph_solver *s, int first, const float *v) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  if (!s->bpos) return fail(WSPH_E_ARG, "call wsph_set_boundary first");
  if (first >= 0 && !v) return fail(WSPH_E_ARG, "null rigid velocity");
  if (first >= 0 && s->nscenes > 0)
    return fail(WSPH_E_UNSUPPORTED, "rigid boundary motion is not available with batched scenes");
  free_rigid(s);
  if (first < 0) return WSPH_OK;
  const int nb = s->nb;
  if (nb > 0) {
    HIP_TRY(dalloc(&s->rb_pos, nb));
    HIP_TRY(dalloc(&s->rb_vel, nb));
    HIP_TRY(dalloc(&s->rb_orig, nb));
    HIP_TRY(dalloc(&s->rb_slot, nb));
    HIP_TRY(dalloc(&s->rb_rank, nb));
    HIP_TRY(dalloc(&s->rb_tmp, nb));
    HIP_TRY(dalloc(&s->rb_perm, nb));
  }
  s->rigid_first = first;
  s->rigid_v_pending = true;
  for (int k = 0; k < 3; ++k) s->rigid_v[k] = v[k];
  return WSPH_OK;
}

// Kernel durations of the calling thread's last wsph_mesh_inside (ms):
// triangle set-up, tile binning, crossing count pass, crossing fill pass,
// classification.
constexpr int kMeshStages = 5;
// then the triangle-tile pairs and the column crossings of that call
static thread_local double g_mesh_ms[kMeshStages + 2] = {};

// Device buffers and events of one wsph_mesh_inside call, released on every
// exit.
struct MeshScratch {
  hipStream_t st = nullptr;
  hipEvent_t ev[2 * kMeshStages] = {};
  MeshTri *tri = nullptr;
  int *bad = nullptr, *ids = nullptr;
  unsigned *cnt = nullptr, *off = nullptr, *fill = nullptr, *tile_cnt = nullptr,
           *tile_off = nullptr, *tile_pos = nullptr;
  int4 *items = nullptr;
  double *z = nullptr;
  ~MeshScratch() {
    if (st) (void)hipStreamSynchronize(st);
    dfree(tri);
    dfree(bad);
    dfree(ids);
    dfree(cnt);
    dfree(off);
    dfree(fill);
    dfree(items);
    dfree(tile_cnt);
    dfree(tile_off);
    dfree(tile_pos);
    dfree(z);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

int wsph_mesh_timings(double *out) {
  if (!out) return fail(WSPH_E_ARG, "null output");
  for (int k = 0; k < kMeshStages + 2; ++k) out[k] = g_mesh_ms[k];
  return WSPH_OK;
}

// Exclusive prefix sum in place; false when the total exceeds 32 bits.
static bool scan_u32(std::vector<unsigned> &v, unsigned long long *total) {
  unsigned long long t = 0;
  for (auto &x : v) {
    const unsigned c = x;
    x = (unsigned)t;
    t += c;
  }
  *total = t;
  return t <= 0xffffffffull;
}

int wsph_mesh_inside(const double *verts, int nv, const int *tris, int nt, const double *ax, int nx,
                     const double *ay, int ny, const double *az, int nz, unsigned char *inside) {
  if (nv < 0 || nt < 0 || nx < 0 || ny < 0 || nz < 0) return fail(WSPH_E_ARG, "negative size");
  const long ncol = (long)nx * ny, npts = ncol * nz;
  for (int k = 0; k < kMeshStages + 2; ++k) g_mesh_ms[k] = 0.0;
  if (npts == 0) return WSPH_OK;
  if (!inside || !ax || !ay || !az) return fail(WSPH_E_ARG, "null lattice or output pointer");
  if (nt > 0 && (!verts || !tris || nv < 1)) return fail(WSPH_E_ARG, "null mesh arrays");
  const long ntiles = mesh_tiles(nx, ny);
  MeshScratch m;
  HIP_TRY(hipStreamCreate(&m.st));
  for (auto &e : m.ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(dalloc(&m.cnt, ncol));
  HIP_TRY(dalloc(&m.off, ncol));
  HIP_TRY(dalloc(&m.fill, ncol));
  HIP_TRY(dalloc(&m.tri, nt));
  HIP_TRY(dalloc(&m.bad, 1));
  HIP_TRY(dalloc(&m.tile_cnt, ntiles));
  HIP_TRY(dalloc(&m.tile_off, ntiles));
  HIP_TRY(dalloc(&m.tile_pos, ntiles));
  HIP_TRY(hipMemsetAsync(m.bad, 0, sizeof(int), m.st));
  HIP_TRY(hipMemsetAsync(m.tile_cnt, 0, sizeof(unsigned) * ntiles, m.st));
  HIP_TRY(hipMemsetAsync(m.tile_pos, 0, sizeof(unsigned) * ntiles, m.st));
  HIP_TRY(hipMemsetAsync(m.cnt, 0, sizeof(unsigned) * ncol, m.st));
  HIP_TRY(hipMemsetAsync(m.fill, 0, sizeof(unsigned) * ncol, m.st));
  launch_mesh_check_axes(ax, nx, m.bad, m.st);
  launch_mesh_check_axes(ay, ny, m.bad, m.st);
  launch_mesh_check_axes(az, nz, m.bad, m.st);
  HIP_TRY(hipEventRecord(m.ev[0], m.st));
  launch_mesh_setup(verts, nv, tris, nt, ax, nx, ay, ny, m.tri, m.tile_cnt, m.bad, m.st);
  HIP_TRY(hipEventRecord(m.ev[1], m.st));
  HIP_TRY(hipGetLastError());
  std::vector<unsigned> h((size_t)ntiles);
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(h.data(), m.tile_cnt, sizeof(unsigned) * ntiles, hipMemcpyDeviceToHost,
                         m.st));
  HIP_TRY(hipMemcpyAsync(&bad, m.bad, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  if (bad & 1) return fail(WSPH_E_ARG, "triangle vertex index outside [0, %d)", nv);
  if (bad & 2) return fail(WSPH_E_ARG, "lattice axes must be strictly ascending");
  const std::vector<unsigned> per_tile = h;
  unsigned long long total = 0;
  if (!scan_u32(h, &total) || total > 0x7fffffffull)
    return fail(WSPH_E_CAPACITY, "%llu triangle-tile pairs", total);
  g_mesh_ms[kMeshStages] = (double)total;
  // work items of the column passes: each tile's triangles in chunks
  const int chunk = mesh_chunk();
  std::vector<int4> items;
  for (long k = 0; k < ntiles; ++k)
    for (unsigned f = 0; f < per_tile[k]; f += chunk)
      items.push_back(make_int4((int)k, (int)(h[k] + f), (int)std::min<unsigned>(chunk, per_tile[k] - f), 0));
  const int nitems = (int)items.size();
  HIP_TRY(dalloc(&m.items, items.size()));
  if (nitems > 0)
    HIP_TRY(hipMemcpyAsync(m.items, items.data(), sizeof(int4) * items.size(),
                           hipMemcpyHostToDevice, m.st));
  HIP_TRY(hipMemcpyAsync(m.tile_off, h.data(), sizeof(unsigned) * ntiles, hipMemcpyHostToDevice,
                         m.st));
  HIP_TRY(dalloc(&m.ids, (size_t)total));
  HIP_TRY(hipEventRecord(m.ev[2], m.st));
  launch_mesh_bin(m.tri, nt, ny, m.tile_off, m.tile_pos, m.ids, m.st);
  HIP_TRY(hipEventRecord(m.ev[3], m.st));
  HIP_TRY(hipEventRecord(m.ev[4], m.st));
  launch_mesh_columns(m.tri, m.ids, m.items, nitems, ax, nx, ay, ny, m.cnt, nullptr, nullptr,
                      nullptr, m.st);
  HIP_TRY(hipEventRecord(m.ev[5], m.st));
  HIP_TRY(hipGetLastError());
  h.resize((size_t)ncol);
  HIP_TRY(hipMemcpyAsync(h.data(), m.cnt, sizeof(unsigned) * ncol, hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  if (!scan_u32(h, &total)) return fail(WSPH_E_CAPACITY, "%llu column crossings", total);
  g_mesh_ms[kMeshStages + 1] = (double)total;
  HIP_TRY(hipMemcpyAsync(m.off, h.data(), sizeof(unsigned) * ncol, hipMemcpyHostToDevice, m.st));
  HIP_TRY(dalloc(&m.z, (size_t)total));
  HIP_TRY(hipEventRecord(m.ev[6], m.st));
  launch_mesh_columns(m.tri, m.ids, m.items, nitems, ax, nx, ay, ny, m.cnt, m.off, m.fill, m.z,
                      m.st);
  HIP_TRY(hipEventRecord(m.ev[7], m.st));
  HIP_TRY(hipEventRecord(m.ev[8], m.st));
  launch_mesh_classify(m.off, m.cnt, m.z, ncol, az, nz, inside, m.st);
  HIP_TRY(hipEventRecord(m.ev[9], m.st));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(m.st));
  for (int k = 0; k < kMeshStages; ++k) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, m.ev[2 * k], m.ev[2 * k + 1]) == hipSuccess) g_mesh_ms[k] = ms;
  }
  return WSPH_OK;
}

int wsph_mesh_inside_noalloc(const double *verts, int nv, const int *tris, int nt, const double *ax,
                             int nx, const double *ay, int ny, const double *az, int nz,
                             unsigned char *inside) {
  if (nv < 0 || nt < 0 || nx < 0 || ny < 0 || nz < 0) return fail(WSPH_E_ARG, "negative size");
  const long ncol = (long)nx * ny, npts = ncol * nz;
  if (npts == 0) return WSPH_OK;
  if (!inside || !ax || !ay || !az) return fail(WSPH_E_ARG, "null lattice or output pointer");
  if (nt > 0 && (!verts || !tris || nv < 1)) return fail(WSPH_E_ARG, "null mesh arrays");
  const long ntiles = mesh_tiles(nx, ny);
  // the part whose size the arguments fix, checked before any device work
  auto fixed = [&](ScratchCarve &c, int **bad, unsigned long long **tot, MeshTri **tri,
                   unsigned **tcnt, unsigned **toff, unsigned **tpos, unsigned **ioff,
                   unsigned **cnt, unsigned **off, unsigned **fill) {
    *bad = c.take<int>(1);
    *tot = c.take<unsigned long long>(2);
    *tri = c.take<MeshTri>((size_t)nt);
    *tcnt = c.take<unsigned>((size_t)ntiles);
    *toff = c.take<unsigned>((size_t)ntiles);
    *tpos = c.take<unsigned>((size_t)ntiles);
    *ioff = c.take<unsigned>((size_t)ntiles);
    *cnt = c.take<unsigned>((size_t)ncol);
    *off = c.take<unsigned>((size_t)ncol);
    *fill = c.take<unsigned>((size_t)ncol);
  };
  int *bad;
  unsigned long long *tot;
  MeshTri *tri;
  unsigned *tcnt, *toff, *tpos, *ioff, *cnt, *off, *fill;
  {
    ScratchCarve sizing{nullptr};
    fixed(sizing, &bad, &tot, &tri, &tcnt, &toff, &tpos, &ioff, &cnt, &off, &fill);
    if (sizing.over)
      return fail(WSPH_E_CAPACITY,
                  "%d triangles on a %d x %d lattice need more than the %zu MiB of mesh scratch",
                  nt, nx, ny, kMeshScratchBytes >> 20);
  }
  unsigned char *base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  ScratchCarve c{base};
  fixed(c, &bad, &tot, &tri, &tcnt, &toff, &tpos, &ioff, &cnt, &off, &fill);
  hipStream_t st = nullptr;  // the null stream: nothing to create
  fill_u32(reinterpret_cast<unsigned *>(bad), 0u, 1, st);
  fill_u32(tcnt, 0u, ntiles, st);
  fill_u32(tpos, 0u, ntiles, st);
  fill_u32(cnt, 0u, ncol, st);
  fill_u32(fill, 0u, ncol, st);
  launch_mesh_check_axes(ax, nx, bad, st);
  launch_mesh_check_axes(ay, ny, bad, st);
  launch_mesh_check_axes(az, nz, bad, st);
  launch_mesh_setup(verts, nv, tris, nt, ax, nx, ay, ny, tri, tcnt, bad, st);
  launch_scan_excl(tcnt, ntiles, 0u, toff, tot, st);
  launch_scan_excl(tcnt, ntiles, (unsigned)mesh_chunk(), ioff, tot + 1, st);
  HIP_TRY(hipGetLastError());
  unsigned long long h_tot[2] = {0, 0};
  int h_bad = 0;
  HIP_TRY(hipMemcpy(h_tot, tot, sizeof(h_tot), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (h_bad & 1) return fail(WSPH_E_ARG, "triangle vertex index outside [0, %d)", nv);
  if (h_bad & 2) return fail(WSPH_E_ARG, "lattice axes must be strictly ascending");
  if (h_tot[0] > 0x7fffffffull) return fail(WSPH_E_CAPACITY, "%llu triangle-tile pairs", h_tot[0]);
  int *ids = c.take<int>((size_t)h_tot[0]);
  int4 *items = c.take<int4>((size_t)h_tot[1]);
  if (c.over)
    return fail(WSPH_E_CAPACITY, "%llu triangle-tile pairs need more than the %zu MiB of mesh scratch",
                h_tot[0], kMeshScratchBytes >> 20);
  const int nitems = (int)h_tot[1];
  launch_mesh_items(tcnt, toff, ioff, ntiles, items, st);
  launch_mesh_bin(tri, nt, ny, toff, tpos, ids, st);
  launch_mesh_columns(tri, ids, items, nitems, ax, nx, ay, ny, cnt, nullptr, nullptr, nullptr, st);
  launch_scan_excl(cnt, ncol, 0u, off, tot, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(h_tot, tot, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  if (h_tot[0] > 0xffffffffull) return fail(WSPH_E_CAPACITY, "%llu column crossings", h_tot[0]);
  double *z = c.take<double>((size_t)h_tot[0]);
  if (c.over)
    return fail(WSPH_E_CAPACITY, "%llu column crossings need more than the %zu MiB of mesh scratch",
                h_tot[0], kMeshScratchBytes >> 20);
  launch_mesh_columns(tri, ids, items, nitems, ax, nx, ay, ny, cnt, off, fill, z, st);
  launch_mesh_classify(off, cnt, z, ncol, az, nz, inside, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return WSPH_OK;
}

int wsph_neighbor_search(wsph_solver *s) {
  if (!s || !s->pos[0]) return fail(WSPH_E_ARG, "no fluid set");
  if (s->timing && s->t_recorded - s->t_processed >= wsph_solver::kRing) timing_fold_one(s);
  launch_reset_state(make_bufs(s), s->stream);
  s->step_count++;
  rec(s, 0);
  int rc = resort_boundary(s);
  if (rc) return rc;
  const int c = s->cur, o = c ^ 1;
  rc = sort_points(s, s->fg, s->pos[c], s->orig[c], s->n, s->slot, s->rank, s->tmp, s->perm,
                       s->sid[c], 0);
  if (rc) return rc;
  gather_fluid(s->perm, s->n, s->pos[c], s->vel[c], s->pprev[c], s->orig[c], s->pos[o], s->vel[o],
               s->pprev[o], s->orig[o], s->stream);
  if (s->nscenes > 0) gather_i32(s->perm, s->n, s->sid[c], s->sid[o], s->stream);
  s->cur = o;
  if (s->P.warm_start == WSPH_WARM_EXTERNAL && s->n > 0)
    k_gather_by_orig<<<nblk(s->n, 256), 256, 0, s->stream>>>(s->p_ext, s->orig[o], s->n,
                                                              s->pstart);
  launch_lists(make_bufs(s), s->C, s->P.block, s->stream);
  HIP_TRY(hipGetLastError());
  s->searched = true;
  return WSPH_OK;
}

int wsph_predict_advection(wsph_solver *s) {
  if (!s || !s->searched) return fail(WSPH_E_ARG, "run wsph_neighbor_search first");
  rec(s, 1);
  const bool learned = s->P.warm_start == WSPH_WARM_LEARNED;
  if (learned && !s->net) return fail(WSPH_E_ARG, "learned start needs wsph_set_predictor");
  if (learned || s->log_pairs) {
    int rc = ensure_learn_bufs(s);
    if (rc) return rc;
  }
  const Bufs b = make_bufs(s);
  const bool plain = s->P.vis_mu == 0.0 && s->P.xsph == 0.0;
  launch_density(b, s->C, plain, s->P.block, s->stream);
  if (!plain) launch_vadv(b, s->C, s->P.block, s->stream);
  launch_rigid_v(s);
  launch_assemble(b, s->C, s->P.warm_start, s->P.block, s->stream);
  if (learned || s->log_pairs) {
    Weights W;
    if (learned) W = weights_of(s->net->dev);
    launch_learned_start(b, s->C, learned ? &W : nullptr, s->stream);
  }
  HIP_TRY(hipGetLastError());
  s->advected = true;
  return WSPH_OK;
}

// Wait until the device has run the stop test of the first `iters` Jacobi
// iterations of step `tag` (the progress word the stop kernel stores in
// pinned memory) or ended the step's loop (done_host == tag): 1 when it
// ended, 0 when it has only progressed, < 0 on a stream error. The host
// spins on the two words; every 1024 polls it also asks the stream, which,
// once idle, has run everything enqueued. No event is recorded: an event
// between two kernels costs ~6 us of device time on this stream.
static int wait_progress(wsph_solver *s, int tag, int iters) {
  const unsigned long long want_tag = (unsigned)tag;
  for (long spin = 1;; ++spin) {
    if (__atomic_load_n(s->done_host, __ATOMIC_ACQUIRE) == tag) return 1;
    const unsigned long long p = __atomic_load_n(s->prog_host, __ATOMIC_ACQUIRE);
    if ((p >> 32) == want_tag && (p & 0xffffffffull) >= (unsigned)iters) return 0;
    if ((spin & 1023) == 0) {
      const hipError_t e = hipStreamQuery(s->stream);
      if (e == hipSuccess) return __atomic_load_n(s->done_host, __ATOMIC_ACQUIRE) == tag ? 1 : 0;
      if (e != hipErrorNotReady) return fail(WSPH_E_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
    }
    __builtin_ia32_pause();
  }
}

// Device-driven Jacobi loop: iterations queued in chunks, one chunk ahead of
// the host's look at the device's progress. `queued` iterations in m0 chunks
// were enqueued already. Before enqueuing chunk m + 1 the host waits for the
// stop tests of chunk m - 1, and it stops enqueuing once the device reports
// the loop ended (chunks already queued return at once).
static int device_loop(wsph_solver *s, const Bufs &b, int src, int queued, int m0) {
  const wsph_params &P = s->P;
  const int C = P.chunk_iters > 0 ? P.chunk_iters : 2;
  const int max_it = P.force_iters > 0 ? P.force_iters : P.max_iters;
  const int tag = s->step_count;
  for (int m = m0; queued < max_it; ++m) {
    const int cnt = C < max_it - queued ? C : max_it - queued;
    hipEvent_t ev3[3] = {nullptr, nullptr, nullptr};
    if (queued == 0) kernel_events(s, ev3);
    launch_jacobi_chunk(b, s->C, P, src, queued, cnt, P.block, s->stream, ev3);
    const int before = queued;  // iterations of the chunks before this one
    queued += cnt;
    if (queued >= max_it) break;
    if (m >= 1) {
      const int rc = wait_progress(s, tag, before);
      if (rc < 0) return rc;
      if (rc == 1) break;
    }
  }
  // every iteration queued: the last one's stop test (the earlier ones ran in
  // the next iteration's J1)
  if (queued >= max_it) launch_jacobi_last_test(b, P, src, max_it - 1, s->stream);
  launch_batch_finalize(b, s->stream);
  s->final_dp = -1;
  return WSPH_OK;
}

int wsph_pressure_solve(wsph_solver *s, wsph_step_stats *stats) {
  if (!s || !s->advected) return fail(WSPH_E_ARG, "run wsph_predict_advection first");
  rec(s, 2);
  const Bufs b = make_bufs(s);
  const wsph_params &P = s->P;
  if (s->n == 0) {  // nothing to solve: a converged step with no iterations
    k_solve_result<<<1, 1, 0, s->stream>>>(s->st, 0, 0.0, 1, 0);
    HIP_TRY(hipGetLastError());
    s->final_dp = 0;
    s->solved = true;
    return WSPH_OK;
  }
  int src = 0;
  if (P.warm_start != WSPH_WARM_ZERO) {
    // A p0 and the first update (IISPH.py:1155-1198, 1209-1219)
    launch_j1(b, s->C, 0, P.block, s->stream);
    launch_j2(b, s->C, 0, -1, P.block, s->stream);
    src = 1;
  }
  if (P.jacobi_mode == WSPH_JACOBI_PERSISTENT) {
    int maxb = 0;
    const int per = persistent_blocks_per_cu(b, s->C, P.block, &maxb);
    if (maxb <= 0 || per <= 0)
      return fail(WSPH_E_UNSUPPORTED, "persistent kernel does not fit on the device");
    // the computed figure, SGPRs included (a launch at the occupancy query's
    // figure, which leaves them out, once found 1792 of 2048 blocks resident)
    int eff = P.coop_blocks_per_cu > 0 ? P.coop_blocks_per_cu : per;
    eff = eff < per ? eff : per;
    const int ncu = maxb / per;
    const int need = (int)nblk(s->n, P.block);
    const int grid = need < eff * ncu ? need : eff * ncu;
    s->persistent_grid = grid;
    if (s->n > 0) {
      hipError_t e = launch_jacobi_persistent(b, s->C, P, src, grid > 0 ? grid : 1, P.block, s->stream);
      if (e != hipSuccess)
        return fail(WSPH_E_UNSUPPORTED, "cooperative launch failed: %s", hipGetErrorString(e));
    }
    s->final_dp = 0;
  } else if (P.jacobi_mode == WSPH_JACOBI_DEVICE) {
    int rc = device_loop(s, b, src, 0, 0);
    if (rc) return rc;
  } else {
    if (s->nscenes > 0)
      return fail(WSPH_E_UNSUPPORTED, "batched scenes need the device or persistent Jacobi driver");
    // one launch per kernel and a host convergence test each iteration,
    // as the reference's while loop (IISPH.py:1200-1296)
    const int max_it = P.force_iters > 0 ? P.force_iters : P.max_iters;
    int k = 0;
    double e = 0.0;
    int cont = 1;
    for (; cont; ++k) {
      launch_j1(b, s->C, src, P.block, s->stream);
      launch_j2(b, s->C, src, k, P.block, s->stream);
      HIP_TRY(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
      cont = jacobi_continue(k, s->st_host->res_sum[k], s->st_host->res_cnt[k], P.min_iters,
                             max_it, P.force_iters, P.eta, &e);
      if (cont) src ^= 1;
    }
    k_solve_result<<<1, 1, 0, s->stream>>>(s->st, k, e, e <= P.eta, k);
    s->final_dp = src;
  }
  HIP_TRY(hipGetLastError());
  s->solved = true;
  (void)stats;
  return WSPH_OK;
}

static int read_stats(wsph_solver *s, wsph_step_stats *stats) {
  HIP_TRY(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const DevState &d = *s->st_host;
  if (stats) {
    stats->iterations = d.iterations;
    stats->converged = d.converged;
    stats->ap_evals = d.ap_evals;
    stats->max_ff_count = d.max_ff;
    stats->max_fs_count = d.max_fs;
    stats->status = d.status;
    stats->residual = d.residual;
    stats->sim_time = d.sim_time;
    stats->v_max = (double)__builtin_bit_cast(float, d.vmax_bits);
  }
  if (d.status & ST_BARRIER) return fail(WSPH_E_BARRIER, "persistent kernel grid barrier timed out");
  if (d.status & (ST_FF_OVERFLOW | ST_FS_OVERFLOW))
    return fail(WSPH_E_CAPACITY, "neighbour list capacity exceeded (fluid %d/%d, boundary %d/%d)",
                d.max_ff, s->P.max_ff, d.max_fs, s->P.max_fs);
  if (d.status & ST_NAN) return fail(WSPH_E_NAN, "non-finite particle state after the step");
  return WSPH_OK;
}

int wsph_integrate(wsph_solver *s, wsph_step_stats *stats) {
  if (!s || !s->solved) return fail(WSPH_E_ARG, "run wsph_pressure_solve first");
  rec(s, 3);
  const Bufs b = make_bufs(s);
  if (stats) {
    // dt used by this step, read before the clock update
    HIP_TRY(hipMemcpyAsync(&s->st_host->dt, &s->st->dt, sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
  }
  // final_dp -1: index kept on the device
  StepEnd end;
  end.dt_mode = s->P.dt_mode;
  end.cfl_d = s->P.cfl * s->P.spacing;
  end.inv_small = s->P.inv_small;
  end.dt_max = s->P.dt_max;
  end.totals = s->totals;
  end.grid = s->persistent_grid;
  // update_rigid_x reads the step's dt, which the clock update below replaces
  launch_rigid_drift(s);
  launch_integrate(b, s->C, s->final_dp, s->P.warm_start, s->P.block, end, s->stream);
  s->have_pairs = s->log_pairs && s->n > 0;
  rec(s, 4);
  HIP_TRY(hipGetLastError());
  s->searched = s->advected = s->solved = false;
  if (s->timing) s->t_recorded++;
  if (stats) {
    double dt_used = 0.0;
    HIP_TRY(hipStreamSynchronize(s->stream));
    dt_used = s->st_host->dt;
    int rc = read_stats(s, stats);
    stats->dt = dt_used;
    return rc;
  }
  return WSPH_OK;
}

// Enqueue the step's front half on the stream (being captured): state reset,
// grid build with buffer set c as the input, neighbour lists, advection
// passes, start value, A p0 and the first graph_iters Jacobi iterations.
static int enqueue_front(wsph_solver *s, int c) {
  const int o = c ^ 1;
  const wsph_params &P = s->P;
  s->cur = c;
  launch_reset_state(make_bufs(s), s->stream);
  int rc = resort_boundary(s);
  if (rc) return rc;
  rc = sort_points(s, s->fg, s->pos[c], s->orig[c], s->n, s->slot, s->rank, s->tmp, s->perm,
                       s->sid[c], 0);
  if (rc) return rc;
  gather_fluid(s->perm, s->n, s->pos[c], s->vel[c], s->pprev[c], s->orig[c], s->pos[o], s->vel[o],
               s->pprev[o], s->orig[o], s->stream);
  if (s->nscenes > 0) gather_i32(s->perm, s->n, s->sid[c], s->sid[o], s->stream);
  s->cur = o;
  if (P.warm_start == WSPH_WARM_EXTERNAL && s->n > 0)
    k_gather_by_orig<<<nblk(s->n, 256), 256, 0, s->stream>>>(s->p_ext, s->orig[o], s->n,
                                                              s->pstart);
  const Bufs b = make_bufs(s);
  launch_lists(b, s->C, P.block, s->stream);
  const bool learned = P.warm_start == WSPH_WARM_LEARNED;
  const bool plain = P.vis_mu == 0.0 && P.xsph == 0.0;
  launch_density(b, s->C, plain, P.block, s->stream);
  if (!plain) launch_vadv(b, s->C, P.block, s->stream);
  launch_assemble(b, s->C, P.warm_start, P.block, s->stream);
  if (learned || s->log_pairs) {
    Weights W;
    if (learned) W = weights_of(s->net->dev);
    launch_learned_start(b, s->C, learned ? &W : nullptr, s->stream);
  }
  int src = 0;
  if (P.warm_start != WSPH_WARM_ZERO) {
    launch_j1(b, s->C, 0, P.block, s->stream);
    launch_j2(b, s->C, 0, -1, P.block, s->stream);
    src = 1;
  }
  s->graph_src = src;
  const int max_it = P.force_iters > 0 ? P.force_iters : P.max_iters;
  const int K = P.graph_iters < max_it ? P.graph_iters : max_it;
  launch_jacobi_chunk(b, s->C, P, src, 0, K, P.block, s->stream);
  HIP_TRY(hipGetLastError());
  return WSPH_OK;
}

static int capture_front(wsph_solver *s, int c) {
  const int saved = s->cur;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_front(s, c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->cur = saved;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(WSPH_E_HIP, "step capture: %s", hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(WSPH_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  s->graph[c] = g;
  s->gexec[c] = x;
  return WSPH_OK;
}

// wsph_step with the front half replayed from a captured graph.
static int graph_step(wsph_solver *s, wsph_step_stats *stats) {
  const wsph_params &P = s->P;
  const bool learned = P.warm_start == WSPH_WARM_LEARNED;
  if (learned && !s->net) return fail(WSPH_E_ARG, "learned start needs wsph_set_predictor");
  if (learned || s->log_pairs) {
    int rc = ensure_learn_bufs(s);  // no allocation may happen inside the capture
    if (rc) return rc;
  }
  const int c = s->cur;
  if (!s->gexec[c]) {
    int rc = capture_front(s, c);
    if (rc) return rc;
  }
  if (s->timing && s->t_recorded - s->t_processed >= wsph_solver::kRing) timing_fold_one(s);
  s->step_count++;
  rec(s, 0);
  HIP_TRY(hipGraphLaunch(s->gexec[c], s->stream));
  s->cur = c ^ 1;
  s->searched = s->advected = true;
  rec(s, 1);
  rec(s, 2);
  const int max_it = P.force_iters > 0 ? P.force_iters : P.max_iters;
  const int K = P.graph_iters < max_it ? P.graph_iters : max_it;
  int rc = device_loop(s, make_bufs(s), s->graph_src, K, 1);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  s->solved = true;
  return wsph_integrate(s, stats);
}

int wsph_step(wsph_solver *s, wsph_step_stats *stats) {
  int rc;
  if (!s) return fail(WSPH_E_ARG, "null solver");
  if (s->P.graph_iters > 0 && s->P.jacobi_mode == WSPH_JACOBI_DEVICE && s->n > 0 &&
      !s->rigid_v_pending)
    return graph_step(s, stats);
  if ((rc = wsph_neighbor_search(s))) return rc;
  if ((rc = wsph_predict_advection(s))) return rc;
  if ((rc = wsph_pressure_solve(s, stats))) return rc;
  return wsph_integrate(s, stats);
}

int wsph_get_field(wsph_solver *s, int field, void *dst) {
  if (!s || !dst) return fail(WSPH_E_ARG, "null argument");
  const int n = s->n;
  const int *orig = s->orig[s->cur];
  const int c = s->cur;
  if (s->final_dp < 0) {
    HIP_TRY(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->final_dp = s->st_host->final_src;
  }
  switch (field) {
    case WSPH_F_X: launch_scatter(s->pos[c], 4, 3, 0, orig, n, dst, s->stream); break;
    case WSPH_F_V: launch_scatter(s->vel[c], 4, 3, 0, orig, n, dst, s->stream); break;
    case WSPH_F_P: launch_scatter(s->pfinal, 1, 1, 0, orig, n, dst, s->stream); break;
    case WSPH_F_RHO: launch_scatter(s->pos[c], 4, 1, 3, orig, n, dst, s->stream); break;
    case WSPH_F_RHO_ADV: launch_scatter(s->aux, 4, 1, 0, orig, n, dst, s->stream); break;
    case WSPH_F_V_ADV: launch_scatter(s->vadv, 4, 3, 0, orig, n, dst, s->stream); break;
    case WSPH_F_D_II: launch_scatter(s->dp[s->final_dp], 4, 3, 0, orig, n, dst, s->stream); break;
    case WSPH_F_A_II: launch_scatter(s->aux, 4, 1, 1, orig, n, dst, s->stream); break;
    case WSPH_F_AP: launch_scatter(s->ap, 1, 1, 0, orig, n, dst, s->stream); break;
    case WSPH_F_S: launch_scatter(s->sap, 4, 3, 0, orig, n, dst, s->stream); break;
    case WSPH_F_FF_COUNT: launch_scatter(s->cnt_ff, 1, 1, 0, orig, n, dst, s->stream); break;
    case WSPH_F_FS_COUNT: launch_scatter(s->cnt_fs, 1, 1, 0, orig, n, dst, s->stream); break;
    case WSPH_F_PSI: launch_scatter(s->bpos, 4, 1, 3, s->borig, s->nb, dst, s->stream); break;
    case WSPH_F_BX: launch_scatter(s->bpos, 4, 3, 0, s->borig, s->nb, dst, s->stream); break;
    case WSPH_F_BV: launch_scatter(s->bvel, 4, 3, 0, s->borig, s->nb, dst, s->stream); break;
    case WSPH_F_CELL_KEY:
      if (n > 0) {
        k_keys4<<<nblk(n, 256), 256, 0, s->stream>>>(s->pos[c], n, s->C.cell_inv, s->P.grid_dims[0],
                                                     s->P.grid_dims[1], s->P.grid_dims[2], s->tmp);
        launch_scatter(s->tmp, 1, 1, 0, orig, n, dst, s->stream);
      }
      break;
    default: return fail(WSPH_E_ARG, "unknown field %d", field);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WSPH_OK;
}

int wsph_get_neighbors(wsph_solver *s, int which, int *cnt, int *idx, int cap) {
  if (!s || !cnt || !idx || cap < 1) return fail(WSPH_E_ARG, "bad arguments");
  if (which != 0 && which != 1) return fail(WSPH_E_ARG, "which must be 0 or 1");
  const int n = s->n;
  if (n > 0) {
    if (which == 0)
      k_export_lists<<<nblk(n, 256), 256, 0, s->stream>>>(s->nbr_ff, s->cnt_ff, n, s->orig[s->cur],
                                                          s->orig[s->cur], s->P.max_ff, cap, cnt, idx);
    else
      k_export_lists<<<nblk(n, 256), 256, 0, s->stream>>>(s->nbr_fs, s->cnt_fs, n, s->orig[s->cur],
                                                          s->borig, s->P.max_fs, cap, cnt, idx);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WSPH_OK;
}

int wsph_get_time(wsph_solver *s, double *sim_time, double *dt) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  HIP_TRY(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (sim_time) *sim_time = s->st_host->sim_time;
  if (dt) *dt = s->st_host->dt;
  return WSPH_OK;
}

int wsph_get_counters(wsph_solver *s, long long *out, int nout) {
  if (!s || !out || nout < 1) return fail(WSPH_E_ARG, "bad arguments");
  long long h[8];
  HIP_TRY(hipMemcpyAsync(h, s->totals, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int k = 0; k < nout && k < 8; ++k) out[k] = h[k];
  return WSPH_OK;
}

int wsph_get_history(wsph_solver *s, wsph_step_record *out, int max_records, int *count) {
  if (!s || !out || max_records < 0) return fail(WSPH_E_ARG, "bad arguments");
  HIP_TRY(hipMemcpyAsync(s->st_host, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int total = s->st_host->step_count;
  const int avail = total < kHistory ? total : kHistory;
  const int m = avail < max_records ? avail : max_records;
  std::vector<StepRecord> h(kHistory);
  HIP_TRY(hipMemcpy(h.data(), s->history, sizeof(StepRecord) * kHistory, hipMemcpyDeviceToHost));
  for (int k = 0; k < m; ++k) {
    const StepRecord &r = h[(total - m + k) % kHistory];
    out[k].step = total - m + k;
    out[k].iterations = r.iterations;
    out[k].converged = r.converged;
    out[k].sim_time = r.sim_time;
    out[k].residual = r.residual;
    out[k].kinetic_energy = r.kinetic;
    out[k].mean_height = r.mean_y;
    out[k].surface_height = r.max_y;
    out[k].v_max = r.v_max;
    out[k].start_source = r.start_source;
    out[k].start_err_learned = r.err_net;
    out[k].start_err_carry = r.err_carry;
  }
  if (count) *count = m;
  return WSPH_OK;
}


int wsph_mlp_create(int max_batch, wsph_mlp **out) {
  if (!out || max_batch < 1) return fail(WSPH_E_ARG, "bad arguments");
  wsph_mlp *m = new wsph_mlp();
  *out = nullptr;
  hipError_t e = mlp_alloc(m->dev, max_batch);
  if (e == hipSuccess) e = hipMalloc((void **)&m->adam_m, sizeof(float) * WSPH_MLP_NPARAM);
  if (e == hipSuccess) e = hipMalloc((void **)&m->adam_v, sizeof(float) * WSPH_MLP_NPARAM);
  if (e == hipSuccess) e = hipMalloc((void **)&m->loss, sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&m->grad, sizeof(float) * WSPH_MLP_NPARAM);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(m->adam_m, 0, sizeof(float) * WSPH_MLP_NPARAM, m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(m->adam_v, 0, sizeof(float) * WSPH_MLP_NPARAM, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    wsph_mlp_destroy(m);
    return fail(WSPH_E_HIP, "predictor allocation: %s", hipGetErrorString(e));
  }
  *out = m;
  return WSPH_OK;
}

int wsph_mlp_destroy(wsph_mlp *m) {
  if (!m) return WSPH_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mlp_free(m->dev);
  for (float *p : {m->adam_m, m->adam_v, m->loss, m->grad})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return WSPH_OK;
}

int wsph_mlp_nparams(void) { return WSPH_MLP_NPARAM; }

int wsph_mlp_set_params(wsph_mlp *m, const float *params) {
  if (!m || !params) return fail(WSPH_E_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(m->dev.params, params, sizeof(float) * WSPH_MLP_NPARAM,
                         hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_m, 0, sizeof(float) * WSPH_MLP_NPARAM, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_v, 0, sizeof(float) * WSPH_MLP_NPARAM, m->stream));
  mlp_sync_weights(m->dev, m->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->step = 0;
  return WSPH_OK;
}

int wsph_mlp_get_params(wsph_mlp *m, float *params) {
  if (!m || !params) return fail(WSPH_E_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(params, m->dev.params, sizeof(float) * WSPH_MLP_NPARAM,
                         hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return WSPH_OK;
}

int wsph_mlp_forward(wsph_mlp *m, const float *x, int n, float *y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return fail(WSPH_E_ARG, "bad arguments");
  mlp_forward(m->dev, x, n, y, m->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(m->stream));
  return WSPH_OK;
}

int wsph_mlp_train_step(wsph_mlp *m, const float *x, const float *t, int n, float lr,
                        float *loss, float *grad) {
  if (!m || n < 1 || !x || !t) return fail(WSPH_E_ARG, "bad arguments");
  if (n > m->dev.max_batch) return fail(WSPH_E_ARG, "batch %d above the maximum %d", n, m->dev.max_batch);
  if (lr > 0.0f) m->step++;
  mlp_train_step(m->dev, m->adam_m, m->adam_v, m->step > 0 ? m->step : 1, lr, x, t, n, m->loss,
                 m->grad, m->stream);
  HIP_TRY(hipGetLastError());
  if (loss) HIP_TRY(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (grad)
    HIP_TRY(hipMemcpyAsync(grad, m->grad, sizeof(float) * WSPH_MLP_NPARAM, hipMemcpyDeviceToHost,
                           m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return WSPH_OK;
}

int wsph_set_scenes(wsph_solver *s, int nscenes, const int *fluid_counts,
                    const int *boundary_counts, const int *grid_dims, const int *bgrid_dims) {
  if (!s || nscenes < 1 || nscenes > kMaxScenes || !fluid_counts || !boundary_counts ||
      !grid_dims || !bgrid_dims)
    return fail(WSPH_E_ARG, "bad scene arguments (1 <= nscenes <= %d)", kMaxScenes);
  if (s->bpos || s->pos[0]) return fail(WSPH_E_ARG, "call wsph_set_scenes before setting particles");
  if (nscenes > 1 && s->P.dt_mode != WSPH_DT_FIXED)
    return fail(WSPH_E_UNSUPPORTED, "batched scenes share one dt: use WSPH_DT_FIXED");
  if (nscenes > 1 && s->P.jacobi_mode == WSPH_JACOBI_HOST)
    return fail(WSPH_E_UNSUPPORTED, "batched scenes need the device or persistent Jacobi driver");
  std::vector<SceneGrid> g(nscenes);
  long fb = 0, bb = 0;
  for (int k = 0; k < nscenes; ++k) {
    if (fluid_counts[k] < 0 || boundary_counts[k] < 0)
      return fail(WSPH_E_ARG, "negative particle count in scene %d", k);
    int rc;
    if ((rc = check_dims(grid_dims + 3 * k, "scene grid")) ||
        (rc = check_dims(bgrid_dims + 3 * k, "scene boundary grid")))
      return rc;
    memcpy(g[k].fdims, grid_dims + 3 * k, sizeof(g[k].fdims));
    memcpy(g[k].bdims, bgrid_dims + 3 * k, sizeof(g[k].bdims));
    g[k].fbase = (int)fb;
    g[k].bbase = (int)bb;
    fb += grid_cells(g[k].fdims, s->P.cell_layout);
    bb += grid_cells(g[k].bdims, s->P.cell_layout);
    if (fb > (1L << 30) || bb > (1L << 30)) return fail(WSPH_E_CAPACITY, "scene grids too large");
  }
  s->nscenes = nscenes;
  s->scene_nf.assign(fluid_counts, fluid_counts + nscenes);
  s->scene_nb.assign(boundary_counts, boundary_counts + nscenes);
  s->scene_grid = g;
  dfree(s->d_scenes);
  HIP_TRY(dalloc(&s->d_scenes, nscenes));
  HIP_TRY(hipMemcpy(s->d_scenes, g.data(), sizeof(SceneGrid) * nscenes, hipMemcpyHostToDevice));
  for (auto p : {&s->sc_done, &s->sc_iters, &s->sc_final}) {
    dfree(*p);
    HIP_TRY(dalloc(p, nscenes));
  }
  dfree(s->sc_res);
  HIP_TRY(dalloc(&s->sc_res, nscenes));
  dfree(s->sc_sum);
  dfree(s->sc_cnt);
  HIP_TRY(dalloc(&s->sc_sum, (size_t)kMaxIters * nscenes));
  HIP_TRY(dalloc(&s->sc_cnt, (size_t)kMaxIters * nscenes));
  int rc = alloc_grid(s->fg, g[0].fdims, s->P.cell_layout, fb);
  if (!rc) rc = alloc_grid(s->bg, g[0].bdims, s->P.cell_layout, bb);
  return rc;
}

int wsph_get_scene_stats(wsph_solver *s, int *iterations, int *converged, double *residual) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  if (s->nscenes == 0) return fail(WSPH_E_ARG, "no batched scenes");
  const int ns = s->nscenes;
  std::vector<int> it(ns);
  std::vector<double> res(ns);
  HIP_TRY(hipMemcpyAsync(it.data(), s->sc_iters, sizeof(int) * ns, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(res.data(), s->sc_res, sizeof(double) * ns, hipMemcpyDeviceToHost,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int k = 0; k < ns; ++k) {
    if (iterations) iterations[k] = it[k];
    if (residual) residual[k] = res[k];
    if (converged) converged[k] = res[k] <= s->P.eta;
  }
  return WSPH_OK;
}

int wsph_set_predictor(wsph_solver *s, wsph_mlp *m) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  if (m != s->net) drop_graphs(s);
  s->net = m;
  return WSPH_OK;
}

int wsph_set_pair_logging(wsph_solver *s, int on) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  if ((on != 0) != s->log_pairs) drop_graphs(s);
  s->log_pairs = on != 0;
  s->have_pairs = false;
  return WSPH_OK;
}

int wsph_get_pairs(wsph_solver *s, float *features, float *targets) {
  if (!s || !features || !targets) return fail(WSPH_E_ARG, "null argument");
  if (!s->have_pairs) return fail(WSPH_E_ARG, "no logged step: enable pair logging and step first");
  const int *orig = s->orig[s->cur];  // the order of the logged step
  launch_scatter(s->feat, WSPH_MLP_IN, WSPH_MLP_IN, 0, orig, s->n, features, s->stream);
  launch_scatter(s->target, 1, 1, 0, orig, s->n, targets, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WSPH_OK;
}

int wsph_hash_keys(const float *x, int n, float cell_width, const int dims[3], int *keys) {
  if (n < 0 || (n > 0 && (!x || !keys)) || !dims || !(cell_width > 0))
    return fail(WSPH_E_ARG, "bad arguments");
  int rc;
  if ((rc = check_dims(dims, "hash"))) return rc;
  linear_keys(x, n, 1.0f / cell_width, dims, keys, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return WSPH_OK;
}

int wsph_set_timing(wsph_solver *s, int enable) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (enable < 0 || enable > 2) return fail(WSPH_E_ARG, "timing mode must be 0, 1 or 2");
  s->timing = enable;
  s->t_recorded = s->t_processed = 0;
  for (double &v : s->t_sum) v = 0.0;
  s->k_sum[0] = s->k_sum[1] = 0.0;
  s->k_steps = 0;
  s->k_pair = 0.0;
  s->k_pairs = 0;
  for (bool &f : s->kev_rec) f = false;
  return WSPH_OK;
}

int wsph_last_timings(wsph_solver *s, double *out, int nout) {
  if (!s || !out) return fail(WSPH_E_ARG, "null argument");
  while (s->t_processed < s->t_recorded) timing_fold_one(s);
  double v[12] = {s->t_sum[0], s->t_sum[1], s->t_sum[2], s->t_sum[3], s->t_sum[4],
                  (double)s->persistent_grid, (double)s->t_processed, s->k_sum[0], s->k_sum[1],
                  (double)s->k_steps, s->k_pair, (double)s->k_pairs};
  for (int k = 0; k < nout && k < 12; ++k) out[k] = v[k];
  return WSPH_OK;
}

int wsph_sync(wsph_solver *s) {
  if (!s) return fail(WSPH_E_ARG, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WSPH_OK;
}

}  // extern "C"
