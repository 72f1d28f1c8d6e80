/*
 * Driver of the CPU oracle (oracle/iisph_oracle.c, oracle/mesh_oracle.c) for
 * the AddressSanitizer / UndefinedBehaviorSanitizer build of
 * tests/test_sanitizers.py: a small dam break with walls run for several
 * steps through every oracle entry point and option (zero and given start,
 * both residual sums, XSPH and the Laplacian term, a forced iteration count,
 * the 101-iteration limit), plus the neighbour, hash-key and mesh
 * containment routines. Exits 0 when every call succeeds.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/wsph.h"
#include "../../oracle/iisph_oracle.h"

int orc_mesh_inside(const double *verts, int nv, const int *tris, int nt, const double *ax, int nx,
                    const double *ay, int ny, const double *az, int nz, unsigned char *inside);

#define CHECK(call)                                              \
  do {                                                           \
    if ((call) < 0) {                                            \
      fprintf(stderr, "%s failed at line %d\n", #call, __LINE__); \
      return 1;                                                  \
    }                                                            \
  } while (0)

static void tables(float h, int nw, float *wt, float *gt) {
  const double sigma = 8.0 / (M_PI * (double)h * h * h);
  for (int k = 0; k < nw; ++k) {
    const double r = (double)h * k / (nw - 1), q = r / h;
    double w = 0.0, g = 0.0;
    if (q <= 0.5) {
      w = sigma * (6 * (q * q * q - q * q) + 1);
      g = (12 - 18 * q) * sigma / ((double)h * h);
    } else if (q < 1.0) {
      w = 2 * sigma * pow(1 - q, 3);
      g = 6 * pow(1 - q, 2) * sigma / (h * r);
    }
    wt[k] = (float)w;
    gt[k] = r <= 1e-10 ? 0.0f : (float)g;
  }
}

int main(void) {
  const float d = 0.02f, h = 0.04f;
  const int nw = 1000, cells = 6, box = 12;
  float *wt = malloc(sizeof(float) * nw), *gt = malloc(sizeof(float) * nw);
  tables(h, nw, wt, gt);
  /* fluid block and a one-layer wall shell around [0, box) */
  const int n = cells * cells * cells;
  float *x = malloc(sizeof(float) * 3 * n), *v = calloc(3 * (size_t)n, sizeof(float));
  for (int i = 0, k = 0; i < cells; ++i)
    for (int j = 0; j < cells; ++j)
      for (int l = 0; l < cells; ++l, ++k) {
        x[3 * k] = (float)(i * (double)d);
        x[3 * k + 1] = (float)(j * (double)d);
        x[3 * k + 2] = (float)(l * (double)d);
      }
  int nb = 0;
  float *xb = malloc(sizeof(float) * 3 * (size_t)(box + 2) * (box + 2) * (box + 2));
  for (int i = -1; i <= box; ++i)
    for (int j = -1; j <= box; ++j)
      for (int l = -1; l <= box; ++l)
        if (i < 0 || j < 0 || l < 0 || i == box || j == box || l == box) {
          xb[3 * nb] = (float)(i * (double)d);
          xb[3 * nb + 1] = (float)(j * (double)d);
          xb[3 * nb + 2] = (float)(l * (double)d);
          ++nb;
        }
  float *vb = calloc(3 * (size_t)nb, sizeof(float)), *psi = malloc(sizeof(float) * nb);
  wsph_params P;
  memset(&P, 0, sizeof P);
  P.spacing = d;
  P.h = h;
  P.rho0 = 1000.0;
  P.mass = 1000.0 * d * d * d;
  P.gravity = 9.81;
  P.omega = 0.5;
  P.eta = 1e-3;
  P.inv_small = 1e-6;
  P.xsph = 0.01;
  P.dt = 1e-3;
  P.dt_max = 1e-3;
  P.cfl = 0.4;
  P.dt_mode = WSPH_DT_FIXED;
  P.min_iters = 2;
  P.max_iters = 101;
  for (int a = 0; a < 3; ++a) P.grid_dims[a] = P.bgrid_dims[a] = 8;
  P.max_ff = 128;
  P.max_fs = 96;
  orc_set_threads(2);
  CHECK(orc_boundary_psi(&P, wt, nw, nb, xb, psi));
  float *f1 = malloc(sizeof(float) * n), *f3 = malloc(sizeof(float) * 3 * n);
  float *rho = malloc(sizeof(float) * n), *va = malloc(sizeof(float) * 3 * n);
  float *ra = malloc(sizeof(float) * n), *dii = malloc(sizeof(float) * 3 * n);
  float *aii = malloc(sizeof(float) * n), *p = malloc(sizeof(float) * n);
  float *ap = malloc(sizeof(float) * n), *pc = malloc(sizeof(float) * n);
  int *cf = malloc(sizeof(int) * n), *cb = malloc(sizeof(int) * n);
  double hist[128];
  orc_out o;
  for (int step = 0; step < 8; ++step) {
    memset(&o, 0, sizeof o);
    if (step != 3) { /* one step without per-particle outputs */
      o.rho = rho, o.v_adv = va, o.rho_adv = ra, o.d_ii = dii, o.a_ii = aii, o.p = p;
      o.Ap = ap, o.s = f3, o.p_carry = pc, o.ff_count = cf, o.fs_count = cb;
    }
    o.res_hist = hist;
    o.res_hist_len = 128;
    P.vis_mu = step == 5 ? 0.5 : 0.0;
    P.force_iters = step == 6 ? 7 : 0;
    P.eta = step == 0 ? 0.0 : 1e-3; /* up to the 101-iteration limit */
    const float *p0 = step >= 2 ? pc : NULL;
    CHECK(orc_step(&P, wt, nw, gt, nw, n, x, v, p0, nb, xb, vb, psi, step & 1, &o));
  }
  /* neighbour lists and keys with aliasing dims */
  const int dims3[3] = {3, 3, 3};
  int *cnt = malloc(sizeof(int) * n), *idx = malloc(sizeof(int) * n * 128);
  float *dist = malloc(sizeof(float) * n * 128);
  CHECK(orc_neighbors(x, n, x, n, h, dims3, 1, 128, cnt, idx, dist));
  CHECK(orc_neighbors(x, n, xb, nb, h, P.bgrid_dims, 0, 128, cnt, idx, dist));
  CHECK(orc_hash_keys(x, n, h, dims3, cf));
  /* mesh containment: an octahedron on a lattice through its vertices */
  const double verts[18] = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  const int tris[24] = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  double ax[9];
  for (int k = 0; k < 9; ++k) ax[k] = -1.0 + 0.25 * k;
  unsigned char inside[9 * 9 * 9];
  CHECK(orc_mesh_inside(verts, 6, tris, 8, ax, 9, ax, 9, ax, 9, inside));
  free(wt), free(gt), free(x), free(v), free(xb), free(vb), free(psi), free(f1), free(f3);
  free(rho), free(va), free(ra), free(dii), free(aii), free(p), free(ap), free(pc), free(cf);
  free(cb), free(cnt), free(idx), free(dist);
  printf("oracle driver ok\n");
  return 0;
}
