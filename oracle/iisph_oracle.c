/*
 * iisph_oracle.c -- CPU float32 restatement of the reference IISPH step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle of libwsph.so.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only to compute expected values or the CPU baseline time; the
 * product path never calls it.
 *
 * Every function restates one piece of the reference (Mo-ZheHan/WarpSPH)
 * and cites it. The reference is NVIDIA-Warp Python; this file follows its
 * arithmetic operation by operation in float32, compiled with
 * -ffp-contract=off so no multiply-add is fused:
 *   - vector expressions evaluate component-wise, left to right;
 *   - x ** 2.0 is evaluated as x * x;
 *   - sums over neighbours run in the hash-grid query order (cells with x
 *     innermost, then y, then z; inside a cell ascending particle index).
 * The warp.HashGrid restated here (cell = trunc(p * (1/w)), + 2^20 origin,
 * clamp at 0, modulo dims, key = cz*dx*dy + cy*dx + cx, stable sort by key,
 * query range [trunc((p-r)/w), min(trunc((p+r)/w), start + dim - 1)]) is
 * third-party behaviour of warp-lang 1.7.1 (not vendored in the reference);
 * its restatement is pinned by brute-force neighbour sets in tests/.
 *
 * One deliberate deviation: the density-error residual of update_rho_error
 * (IISPH.py:542-568) is summed exactly in 64-bit fixed point (each term
 * truncated to a multiple of 2^-28 and clamped at 1024) so that the result
 * does not depend on summation order; the reference's float32 atomic sum
 * has no fixed order on a GPU. residual_mode = 1 selects the sequential
 * float32 sum of a one-thread run instead.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/wsph.h"
#include "iisph_oracle.h"

/* ------------------------------------------------------------------------ */
/* warp.HashGrid restatement (used at IISPH.py:771-780, 811, 1414)           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int dx, dy, dz;
  float inv;
  int ncell;
  int *start, *end; /* per hashed cell */
  int *ids;         /* point ids in cell order */
} ogrid;

static int ogrid_index(const ogrid *g, int x, int y, int z) {
  const int origin = 1 << 20;
  x += origin;
  y += origin;
  z += origin;
  if (x < 0) x = 0;
  if (y < 0) y = 0;
  if (z < 0) z = 0;
  const int cx = x % g->dx, cy = y % g->dy, cz = z % g->dz;
  return cz * (g->dx * g->dy) + cy * g->dx + cx;
}

static int ogrid_key(const ogrid *g, const float *p) {
  return ogrid_index(g, (int)(p[0] * g->inv), (int)(p[1] * g->inv), (int)(p[2] * g->inv));
}

static int ogrid_build(ogrid *g, const float *pts, int n, float w, const int dims[3]) {
  g->dx = dims[0];
  g->dy = dims[1];
  g->dz = dims[2];
  g->inv = 1.0f / w;
  g->ncell = g->dx * g->dy * g->dz;
  g->start = (int *)calloc((size_t)g->ncell + 1, sizeof(int));
  g->end = (int *)calloc((size_t)g->ncell, sizeof(int));
  g->ids = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *key = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!g->start || !g->end || !g->ids || !key) return -1;
  for (int i = 0; i < n; ++i) {
    key[i] = ogrid_key(g, pts + 3 * (size_t)i);
    g->start[key[i] + 1]++;
  }
  for (int c = 0; c < g->ncell; ++c) g->start[c + 1] += g->start[c];
  for (int c = 0; c < g->ncell; ++c) g->end[c] = g->start[c];
  /* stable placement: ascending id inside each cell */
  for (int i = 0; i < n; ++i) g->ids[g->end[key[i]]++] = i;
  free(key);
  return 0;
}

static void ogrid_free(ogrid *g) {
  free(g->start);
  free(g->end);
  free(g->ids);
  memset(g, 0, sizeof(*g));
}

typedef struct {
  int xs, ys, zs, xe, ye, ze, x, y, z, ci, ce;
} oquery;

static void oquery_begin(oquery *q, const ogrid *g, const float *p, float r) {
  q->xs = (int)((p[0] - r) * g->inv);
  q->ys = (int)((p[1] - r) * g->inv);
  q->zs = (int)((p[2] - r) * g->inv);
  int xe = (int)((p[0] + r) * g->inv), ye = (int)((p[1] + r) * g->inv),
      ze = (int)((p[2] + r) * g->inv);
  q->xe = xe < q->xs + g->dx - 1 ? xe : q->xs + g->dx - 1;
  q->ye = ye < q->ys + g->dy - 1 ? ye : q->ys + g->dy - 1;
  q->ze = ze < q->zs + g->dz - 1 ? ze : q->zs + g->dz - 1;
  q->x = q->xs;
  q->y = q->ys;
  q->z = q->zs;
  const int c = ogrid_index(g, q->x, q->y, q->z);
  q->ci = g->start[c];
  q->ce = g->end[c];
}

static int oquery_next(oquery *q, const ogrid *g, int *idx) {
  for (;;) {
    if (q->ci < q->ce) {
      *idx = g->ids[q->ci++];
      return 1;
    }
    q->x++;
    if (q->x > q->xe) {
      q->x = q->xs;
      q->y++;
    }
    if (q->y > q->ye) {
      q->y = q->ys;
      q->z++;
    }
    if (q->z > q->ze) return 0;
    const int c = ogrid_index(g, q->x, q->y, q->z);
    q->ci = g->start[c];
    q->ce = g->end[c];
  }
}

/* wp.length(a - b) = sqrt(dot(d, d)), dot summed left to right */
static inline float dist3(const float *a, const float *b) {
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* ------------------------------------------------------------------------ */
/* Kernel tables: spline_W / grad_spline_W (sph_funcs.py:92-118)            */
/* ------------------------------------------------------------------------ */
typedef struct {
  const float *wt, *gt;
  int nw, ng;
  float h, w_scale, g_scale;
} otab;

static void otab_init(otab *t, const float *wt, int nw, const float *gt, int ng, double h) {
  t->wt = wt;
  t->gt = gt;
  t->nw = nw;
  t->ng = ng;
  t->h = (float)h;
  const float h_inv = (float)(1.0 / h);
  /* float(W_SIZE - 1) * h_inv, float32 (sph_funcs.py:98, 113) */
  t->w_scale = (float)(nw - 1) * h_inv;
  t->g_scale = (float)(ng - 1) * h_inv;
}

/* spline_W, sph_funcs.py:92-102 */
static inline float spline_W(const otab *t, float r) {
  if (r >= t->h) return 0.0f;
  const float idx_float = r * t->w_scale;
  int lo = (int)idx_float;
  if (lo > t->nw - 2) lo = t->nw - 2;
  const float alpha = idx_float - (float)lo;
  return (1.0f - alpha) * t->wt[lo] + alpha * t->wt[lo + 1];
}

/* grad_spline_W, sph_funcs.py:105-118; r is x_j - x_i */
static inline void grad_spline_W(const otab *t, const float *r, float *out) {
  const float r_len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (r_len >= t->h || r_len < 1e-10f) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float idx_float = r_len * t->g_scale;
  int lo = (int)idx_float;
  if (lo > t->ng - 2) lo = t->ng - 2;
  const float alpha = idx_float - (float)lo;
  const float scalar = (1.0f - alpha) * t->gt[lo] + alpha * t->gt[lo + 1];
  out[0] = scalar * r[0];
  out[1] = scalar * r[1];
  out[2] = scalar * r[2];
}

static inline float dot3(const float *a, const float *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ------------------------------------------------------------------------ */
/* CSR neighbour lists (count_/store_same_neighbor IISPH.py:52-104,          */
/* count_/store_diff_neighbor IISPH.py:107-163)                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  int *num, *off, *list;
  float *dist;
  int total, max_count;
} ocsr;

static int build_csr(ocsr *c, const float *xq, int nq, const float *xs, const ogrid *g,
                     float h, int exclude_self) {
  c->num = (int *)calloc((size_t)(nq > 0 ? nq : 1), sizeof(int));
  c->off = (int *)calloc((size_t)nq + 1, sizeof(int));
  if (!c->num || !c->off) return -1;
  c->max_count = 0;
  /* count pass */
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nq; ++i) {
    const float *x = xq + 3 * (size_t)i;
    oquery q;
    int idx, cnt = 0;
    oquery_begin(&q, g, x, h);
    while (oquery_next(&q, g, &idx)) {
      const float d = dist3(x, xs + 3 * (size_t)idx);
      if (d < h && (!exclude_self || idx != i)) cnt++;
    }
    c->num[i] = cnt;
  }
  for (int i = 0; i < nq; ++i) {
    c->off[i + 1] = c->off[i] + c->num[i];
    if (c->num[i] > c->max_count) c->max_count = c->num[i];
  }
  c->total = c->off[nq];
  c->list = (int *)malloc(sizeof(int) * (size_t)(c->total > 0 ? c->total : 1));
  c->dist = (float *)malloc(sizeof(float) * (size_t)(c->total > 0 ? c->total : 1));
  if (!c->list || !c->dist) return -1;
  /* store pass */
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nq; ++i) {
    const float *x = xq + 3 * (size_t)i;
    oquery q;
    int idx, k = c->off[i];
    oquery_begin(&q, g, x, h);
    while (oquery_next(&q, g, &idx)) {
      const float d = dist3(x, xs + 3 * (size_t)idx);
      if (d < h && (!exclude_self || idx != i)) {
        c->list[k] = idx;
        c->dist[k] = d;
        k++;
      }
    }
  }
  return 0;
}

static void free_csr(ocsr *c) {
  free(c->num);
  free(c->off);
  free(c->list);
  free(c->dist);
  memset(c, 0, sizeof(*c));
}

/* ------------------------------------------------------------------------ */
/* Public test entry points                                                  */
/* ------------------------------------------------------------------------ */
void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
}

int orc_hash_keys(const float *x, int n, float cell_width, const int dims[3], int *keys) {
  ogrid g;
  memset(&g, 0, sizeof(g));
  g.dx = dims[0];
  g.dy = dims[1];
  g.dz = dims[2];
  g.inv = 1.0f / cell_width;
  for (int i = 0; i < n; ++i) keys[i] = ogrid_key(&g, x + 3 * (size_t)i);
  return 0;
}

int orc_neighbors(const float *xq, int nq, const float *xs, int ns, float h, const int dims[3],
                  int exclude_self, int cap, int *cnt, int *idx, float *dist) {
  ogrid g;
  ocsr c;
  memset(&c, 0, sizeof(c));
  if (ogrid_build(&g, xs, ns, h, dims)) return -1;
  if (build_csr(&c, xq, nq, xs, &g, h, exclude_self)) return -1;
  for (int i = 0; i < nq; ++i) {
    cnt[i] = c.num[i];
    for (int k = 0; k < c.num[i] && k < cap; ++k) {
      idx[(size_t)i * cap + k] = c.list[c.off[i] + k];
      if (dist) dist[(size_t)i * cap + k] = c.dist[c.off[i] + k];
    }
  }
  const int m = c.max_count;
  free_csr(&c);
  ogrid_free(&g);
  return m;
}

/* compute_boundary_density, IISPH.py:23-47 */
int orc_boundary_psi(const wsph_params *P, const float *wt, int nw, int nb, const float *xb,
                     float *psi) {
  otab t;
  otab_init(&t, wt, nw, wt, nw, P->h);
  const float h = (float)P->h, rho0 = (float)P->rho0;
  ogrid g;
  if (ogrid_build(&g, xb, nb, h, P->bgrid_dims)) return -1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nb; ++i) {
    const float *x = xb + 3 * (size_t)i;
    float rho = spline_W(&t, 0.0f);
    oquery q;
    int idx;
    oquery_begin(&q, &g, x, h);
    while (oquery_next(&q, &g, &idx)) {
      const float d = dist3(x, xb + 3 * (size_t)idx);
      if (d < h && idx != i) rho += spline_W(&t, d);
    }
    psi[i] = rho0 / rho;
  }
  ogrid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* One IISPH step, IISPH.step (IISPH.py:1002-1371)                           */
/* ------------------------------------------------------------------------ */
int orc_step(const wsph_params *P, const float *wt, int nw, const float *gt, int ng, int n,
             float *x, float *v, const float *p0, int nb, const float *xb, const float *vb,
             const float *psi, int residual_mode, orc_out *out) {
  otab t;
  otab_init(&t, wt, nw, gt, ng, P->h);
  const float h = (float)P->h;
  const float m = (float)P->mass;
  const float rho0 = (float)P->rho0;
  const float omega = (float)P->omega;
  const float inv_small = (float)P->inv_small;
  const float g_y = -(float)P->gravity;
  const float vis_mu = (float)P->vis_mu;
  const float xsph = (float)P->xsph;
  const float dt = (float)P->dt;
  const float inv_dt = (float)(1.0 / P->dt);
  const int nn = n > 0 ? n : 1;

  /* neighbor_search, IISPH.py:1408-1482 */
  ogrid gf, gb;
  ocsr ff, fs;
  memset(&ff, 0, sizeof(ff));
  memset(&fs, 0, sizeof(fs));
  if (ogrid_build(&gf, x, n, h, P->grid_dims)) return -1;
  if (ogrid_build(&gb, xb, nb, h, P->bgrid_dims)) return -1;
  if (build_csr(&ff, x, n, x, &gf, h, 1)) return -1;
  if (build_csr(&fs, x, n, xb, &gb, h, 0)) return -1;

  float *rho = (float *)malloc(sizeof(float) * nn);
  float *v_adv = (float *)malloc(sizeof(float) * 3 * nn);
  float *rho_adv = (float *)malloc(sizeof(float) * nn);
  float *d_ii = (float *)malloc(sizeof(float) * 3 * nn);
  float *a_ii = (float *)malloc(sizeof(float) * nn);
  float *p = (float *)malloc(sizeof(float) * nn);
  float *Ap = (float *)malloc(sizeof(float) * nn);
  float *s = (float *)malloc(sizeof(float) * 3 * nn);
  const int np = ff.total > 0 ? ff.total : 1;
  float *d_ij = (float *)malloc(sizeof(float) * 3 * np);
  float *d_ji = (float *)malloc(sizeof(float) * 3 * np);
  if (!rho || !v_adv || !rho_adv || !d_ii || !a_ii || !p || !Ap || !s || !d_ij || !d_ji)
    return -1;

  /* init_pressure, IISPH.py:166-174 (p = 0), or the caller's start value */
  for (int i = 0; i < n; ++i) p[i] = p0 ? p0[i] : 0.0f;

  /* compute_density, IISPH.py:177-214 */
  const float W0 = spline_W(&t, 0.0f);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    float t1 = 0.0f, t2 = 0.0f;
    for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) t1 += spline_W(&t, ff.dist[j]);
    for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j)
      t2 += spline_W(&t, fs.dist[j]) * psi[fs.list[j]];
    rho[i] = t1 * m + t2 + W0 * m;
  }

  /* predict_v_adv, IISPH.py:217-278; plus XSPH smoothing when xsph != 0 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = x + 3 * (size_t)i, *vi = v + 3 * (size_t)i;
    float lap[3] = {0.0f, 0.0f, 0.0f};
    if (vis_mu != 0.0f) {
      float t1[3] = {0.0f, 0.0f, 0.0f}, t2[3] = {0.0f, 0.0f, 0.0f};
      for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {
        const int k = ff.list[j];
        const float *xk = x + 3 * (size_t)k, *vk = v + 3 * (size_t)k;
        const float xij[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};
        float gw[3];
        grad_spline_W(&t, xij, gw);
        const float vd[3] = {vi[0] - vk[0], vi[1] - vk[1], vi[2] - vk[2]};
        const float dd = dot3(vd, xij);
        const float r2 = ff.dist[j] * ff.dist[j];
        for (int c = 0; c < 3; ++c) t1[c] = t1[c] - gw[c] * dd / rho[k] / r2;
      }
      for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j) {
        const int b = fs.list[j];
        const float *xbb = xb + 3 * (size_t)b;
        const float xib[3] = {xbb[0] - xi[0], xbb[1] - xi[1], xbb[2] - xi[2]};
        float gw[3];
        grad_spline_W(&t, xib, gw);
        const float dd = dot3(vi, xib);
        const float r2 = fs.dist[j] * fs.dist[j];
        for (int c = 0; c < 3; ++c) t2[c] = t2[c] - gw[c] * psi[b] * dd / r2;
      }
      for (int c = 0; c < 3; ++c) lap[c] = 10.0f * (t1[c] * m + t2[c] / rho0);
    }
    const float gv[3] = {0.0f, g_y, 0.0f};
    float va[3];
    for (int c = 0; c < 3; ++c) va[c] = vi[c] + (lap[c] * vis_mu / rho[i] + gv[c]) * dt;
    if (xsph != 0.0f) {
      /* XSPH: v_adv += c * sum_j (v_j - v_i) * ((m / rho_j) * W_ij) */
      float acc[3] = {0.0f, 0.0f, 0.0f};
      for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {
        const int k = ff.list[j];
        const float *vk = v + 3 * (size_t)k;
        const float coef = m / rho[k] * spline_W(&t, ff.dist[j]);
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + (vk[c] - vi[c]) * coef;
      }
      for (int c = 0; c < 3; ++c) va[c] = va[c] + acc[c] * xsph;
    }
    for (int c = 0; c < 3; ++c) v_adv[3 * (size_t)i + c] = va[c];
  }

  /* predict_rho_adv, IISPH.py:281-329 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = x + 3 * (size_t)i, *va = v_adv + 3 * (size_t)i;
    float t1 = 0.0f, t2 = 0.0f;
    for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {
      const int k = ff.list[j];
      const float *xk = x + 3 * (size_t)k, *vk = v_adv + 3 * (size_t)k;
      const float vij[3] = {va[0] - vk[0], va[1] - vk[1], va[2] - vk[2]};
      const float r[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      t1 += dot3(vij, gw);
    }
    for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j) {
      const int b = fs.list[j];
      const float *xbb = xb + 3 * (size_t)b;
      const float vbb[3] = {vb ? vb[3 * (size_t)b] : 0.0f, vb ? vb[3 * (size_t)b + 1] : 0.0f,
                            vb ? vb[3 * (size_t)b + 2] : 0.0f};
      const float vib[3] = {va[0] - vbb[0], va[1] - vbb[1], va[2] - vbb[2]};
      const float r[3] = {xbb[0] - xi[0], xbb[1] - xi[1], xbb[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      t2 += dot3(vib, gw) * psi[b];
    }
    rho_adv[i] = rho[i] + (t1 * m + t2) * dt;
  }

  /* compute_term_d, IISPH.py:332-382 */
  const float dt2 = dt * dt;
  const float term_3 = m * dt2;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = x + 3 * (size_t)i;
    const float ri = rho[i];
    const float ri2 = ri * ri;
    float t1[3] = {0.0f, 0.0f, 0.0f}, t2[3] = {0.0f, 0.0f, 0.0f};
    for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {
      const int k = ff.list[j];
      const float *xk = x + 3 * (size_t)k;
      const float r[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      const float rk2 = rho[k] * rho[k];
      for (int c = 0; c < 3; ++c) {
        t1[c] = t1[c] - gw[c];
        d_ij[3 * (size_t)j + c] = -gw[c] * term_3 / rk2;
        d_ji[3 * (size_t)j + c] = gw[c] * term_3 / ri2;
      }
    }
    for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j) {
      const int b = fs.list[j];
      const float *xbb = xb + 3 * (size_t)b;
      const float r[3] = {xbb[0] - xi[0], xbb[1] - xi[1], xbb[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      for (int c = 0; c < 3; ++c) t2[c] = t2[c] - gw[c] * psi[b];
    }
    for (int c = 0; c < 3; ++c) d_ii[3 * (size_t)i + c] = (t1[c] * m + t2[c]) * dt2 / ri2;
  }

  /* compute_term_a, IISPH.py:385-428 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = x + 3 * (size_t)i, *di = d_ii + 3 * (size_t)i;
    float t1 = 0.0f, t2 = 0.0f;
    for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {
      const int k = ff.list[j];
      const float *xk = x + 3 * (size_t)k;
      const float r[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      const float dd[3] = {di[0] - d_ji[3 * (size_t)j], di[1] - d_ji[3 * (size_t)j + 1],
                           di[2] - d_ji[3 * (size_t)j + 2]};
      t1 += dot3(dd, gw);
    }
    for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j) {
      const int b = fs.list[j];
      const float *xbb = xb + 3 * (size_t)b;
      const float r[3] = {xbb[0] - xi[0], xbb[1] - xi[1], xbb[2] - xi[2]};
      float gw[3];
      grad_spline_W(&t, r, gw);
      t2 += dot3(di, gw) * psi[b];
    }
    a_ii[i] = t1 * m + t2;
  }

  /* compute_term_Ap_1 (IISPH.py:431-466) and compute_term_Ap_2 (IISPH.py:469-518) */
#define ORC_AP()                                                                          \
  do {                                                                                    \
    _Pragma("omp parallel for schedule(static)") for (int i = 0; i < n; ++i) {            \
      const float *xi = x + 3 * (size_t)i;                                                \
      float t1 = 0.0f;                                                                    \
      float si[3] = {0.0f, 0.0f, 0.0f};                                                   \
      for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {                           \
        const int k = ff.list[j];                                                         \
        const float *xk = x + 3 * (size_t)k;                                              \
        const float term[3] = {d_ji[3 * (size_t)j] * p[i] - d_ii[3 * (size_t)k] * p[k],   \
                               d_ji[3 * (size_t)j + 1] * p[i] - d_ii[3 * (size_t)k + 1] * p[k], \
                               d_ji[3 * (size_t)j + 2] * p[i] - d_ii[3 * (size_t)k + 2] * p[k]}; \
        const float r[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};                 \
        float gw[3];                                                                      \
        grad_spline_W(&t, r, gw);                                                         \
        t1 += dot3(term, gw);                                                             \
        for (int c = 0; c < 3; ++c) si[c] = si[c] + d_ij[3 * (size_t)j + c] * p[k];       \
      }                                                                                   \
      for (int c = 0; c < 3; ++c) s[3 * (size_t)i + c] = si[c];                           \
      Ap[i] = t1 * m;                                                                     \
    }                                                                                     \
    _Pragma("omp parallel for schedule(static)") for (int i = 0; i < n; ++i) {            \
      const float *xi = x + 3 * (size_t)i, *si = s + 3 * (size_t)i;                       \
      float t1 = 0.0f, t2 = 0.0f;                                                         \
      for (int j = ff.off[i]; j < ff.off[i] + ff.num[i]; ++j) {                           \
        const int k = ff.list[j];                                                         \
        const float *xk = x + 3 * (size_t)k, *sk = s + 3 * (size_t)k;                     \
        const float r[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};                 \
        const float sd[3] = {si[0] - sk[0], si[1] - sk[1], si[2] - sk[2]};                \
        float gw[3];                                                                      \
        grad_spline_W(&t, r, gw);                                                         \
        t1 += dot3(sd, gw);                                                               \
      }                                                                                   \
      for (int j = fs.off[i]; j < fs.off[i] + fs.num[i]; ++j) {                           \
        const int b = fs.list[j];                                                         \
        const float *xbb = xb + 3 * (size_t)b;                                            \
        const float r[3] = {xbb[0] - xi[0], xbb[1] - xi[1], xbb[2] - xi[2]};              \
        float gw[3];                                                                      \
        grad_spline_W(&t, r, gw);                                                         \
        t2 += dot3(si, gw) * psi[b];                                                      \
      }                                                                                   \
      Ap[i] = Ap[i] + (t1 * m + t2 + a_ii[i] * p[i]);                                     \
    }                                                                                     \
  } while (0)

  ORC_AP();
  /* A p of a zero start is identically zero; count only evaluations that a
   * zero start cannot skip */
  int ap_evals = p0 ? 1 : 0;

  /* Jacobi loop, IISPH.py:1200-1296: at least min_iters, at most max_iters,
   * stop once the mean density error is <= eta */
  const int max_it = P->force_iters > 0 ? P->force_iters : P->max_iters;
  int it = 0;
  double e = 0.0;
  for (; it < max_it;) {
    /* update_p, IISPH.py:521-539 */
    for (int i = 0; i < n; ++i) {
      float a = a_ii[i];
      if (a > -inv_small && a < inv_small) a = (a < 0.0f ? -1.0f : 1.0f) * inv_small;
      float pi = p[i] + (rho0 - rho_adv[i] - Ap[i]) * omega / a;
      p[i] = pi > 0.0f ? pi : 0.0f;
    }
    ORC_AP();
    ap_evals++;
    /* update_rho_error, IISPH.py:542-568, and average_rho_error, IISPH.py:1489-1494 */
    uint64_t sum_fx = 0;
    float sum_f = 0.0f;
    long cnt = 0;
    for (int i = 0; i < n; ++i) {
      const float ur = rho_adv[i] + Ap[i];
      if (ur > rho0) {
        float term = fabsf(ur / rho0 - 1.0f);
        sum_f += term;
        if (term > WSPH_RESIDUAL_TERM_CLAMP) term = WSPH_RESIDUAL_TERM_CLAMP;
        sum_fx += (uint64_t)((double)term * WSPH_RESIDUAL_FIXED_SCALE);
        cnt++;
      }
    }
    if (cnt == 0)
      e = 0.0;
    else if (residual_mode == 1)
      e = (double)sum_f / (double)cnt;
    else
      e = ((double)sum_fx / WSPH_RESIDUAL_FIXED_SCALE) / (double)cnt;
    if (out && out->res_hist && it < out->res_hist_len) out->res_hist[it] = e;
    it++;
    if (P->force_iters > 0) continue;
    if (it >= P->min_iters && e <= P->eta) break;
  }
#undef ORC_AP

  /* kick, IISPH.py:674-688 and drift, IISPH.py:691-699 */
  float vmax = 0.0f;
  for (int i = 0; i < n; ++i) {
    float vn[3];
    for (int c = 0; c < 3; ++c) {
      const float dv = inv_dt * (d_ii[3 * (size_t)i + c] * p[i] + s[3 * (size_t)i + c]);
      vn[c] = v_adv[3 * (size_t)i + c] + dv;
      v[3 * (size_t)i + c] = vn[c];
    }
    const float l = sqrtf(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    if (l > vmax) vmax = l;
  }
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) x[3 * (size_t)i + c] = x[3 * (size_t)i + c] + v[3 * (size_t)i + c] * dt;

  if (out) {
    out->iterations = it;
    out->ap_evals = ap_evals;
    out->residual = e;
    out->converged = e <= P->eta;
    out->v_max = vmax;
    out->max_ff = ff.max_count;
    out->max_fs = fs.max_count;
#define ORC_COPY(field, src, k)                                 \
  if (out->field) memcpy(out->field, src, sizeof(float) * (k) * (size_t)n)
    ORC_COPY(rho, rho, 1);
    ORC_COPY(v_adv, v_adv, 3);
    ORC_COPY(rho_adv, rho_adv, 1);
    ORC_COPY(d_ii, d_ii, 3);
    ORC_COPY(a_ii, a_ii, 1);
    ORC_COPY(p, p, 1);
    ORC_COPY(Ap, Ap, 1);
    ORC_COPY(s, s, 3);
#undef ORC_COPY
    if (out->p_carry)
      for (int i = 0; i < n; ++i)
        out->p_carry[i] = (a_ii[i] > -inv_small && a_ii[i] < inv_small) ? 0.0f : p[i];
    if (out->ff_count) memcpy(out->ff_count, ff.num, sizeof(int) * (size_t)n);
    if (out->fs_count) memcpy(out->fs_count, fs.num, sizeof(int) * (size_t)n);
  }

  free(rho);
  free(v_adv);
  free(rho_adv);
  free(d_ii);
  free(a_ii);
  free(p);
  free(Ap);
  free(s);
  free(d_ij);
  free(d_ji);
  free_csr(&ff);
  free_csr(&fs);
  ogrid_free(&gf);
  ogrid_free(&gb);
  return 0;
}
