/*
 * iisph_oracle.h -- CPU float32 restatement of the reference IISPH step.
 * TEST INFRASTRUCTURE ONLY (the oracle of libwsph.so); see iisph_oracle.c.
 */
#ifndef IISPH_ORACLE_H
#define IISPH_ORACLE_H

#include "../include/wsph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_out {
  /* optional per-particle outputs (NULL to skip), original order */
  float *rho, *v_adv, *rho_adv, *d_ii, *a_ii, *p, *Ap, *s;
  float *p_carry; /* p, zero where |a_ii| < inv_small (start value of the next step) */
  int *ff_count, *fs_count;
  double *res_hist; /* residual after each Jacobi iteration */
  int res_hist_len;
  /* scalars */
  int iterations, ap_evals, converged, max_ff, max_fs;
  double residual;
  float v_max;
} orc_out;

void orc_set_threads(int nthreads);
int orc_hash_keys(const float *x, int n, float cell_width, const int dims[3], int *keys);
int orc_neighbors(const float *xq, int nq, const float *xs, int ns, float h, const int dims[3],
                  int exclude_self, int cap, int *cnt, int *idx, float *dist);
int orc_boundary_psi(const wsph_params *P, const float *wt, int nw, int nb, const float *xb,
                     float *psi);
int orc_step(const wsph_params *P, const float *wt, int nw, const float *gt, int ng, int n,
             float *x, float *v, const float *p0, int nb, const float *xb, const float *vb,
             const float *psi, int residual_mode, orc_out *out);

#ifdef __cplusplus
}
#endif
#endif
