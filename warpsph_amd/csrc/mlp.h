// mlp.h -- per-particle pressure predictor (8 -> 64 -> 64 -> 1, SiLU).
#pragma once

#include <hip/hip_runtime.h>

#define WSPH_MLP_IN 8
#define WSPH_MLP_H 64
// parameter vector (fp32): W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3
#define WSPH_MLP_OFF_W1 0
#define WSPH_MLP_OFF_B1 (WSPH_MLP_OFF_W1 + WSPH_MLP_H * WSPH_MLP_IN)
#define WSPH_MLP_OFF_W2 (WSPH_MLP_OFF_B1 + WSPH_MLP_H)
#define WSPH_MLP_OFF_B2 (WSPH_MLP_OFF_W2 + WSPH_MLP_H * WSPH_MLP_H)
#define WSPH_MLP_OFF_W3 (WSPH_MLP_OFF_B2 + WSPH_MLP_H)
#define WSPH_MLP_OFF_B3 (WSPH_MLP_OFF_W3 + WSPH_MLP_H)
#define WSPH_MLP_NPARAM (WSPH_MLP_OFF_B3 + 1)

namespace wsph {

struct MlpDev {
  float *params;                 // fp32 master parameters
  __bf16 *w1, *w2, *w3;          // bf16 copies fed to the matrix cores
  float *part, *sse_part;        // per-block gradient partials
  int nparts, max_batch;
};

hipError_t mlp_alloc(MlpDev &m, int max_batch);
void mlp_free(MlpDev &m);
void mlp_sync_weights(const MlpDev &m, hipStream_t s);
void mlp_forward(const MlpDev &m, const float *x, int n, float *y, hipStream_t s);
void mlp_train_step(MlpDev &m, float *adam_m, float *adam_v, int step, float lr, const float *x,
                    const float *t, int n, float *loss, float *grad, hipStream_t s);

}  // namespace wsph
