// Launchers of the learned-weno5 kernels (lsg_learned.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsg {

constexpr int MLP_PSTRIDE = 1348;  // partial-gradient row: 1347 parameters + squared-residual sum

// Gradient of the mean squared loss over samples X[n][8], T[n] at the fp32
// parameters P: per-workgroup partials (rows x MLP_PSTRIDE floats, rows >= 1)
// summed into grad[MLP_PSTRIDE] (grad[1347] = sum of squared residuals).
void launchMlpGrad(hipStream_t s, const float *X, const float *T, uint32_t n, const float *P, float *partials,
                   int rows, float *grad);
// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at step t >= 1.
void launchMlpAdam(hipStream_t s, float *P, float *M, float *V, const float *G, int t, float lr);
// Learned weno5 on n stencils (x[n*7]) with bf16 parameters.
void launchWenoLearned(hipStream_t s, uint32_t n, const double *x, double dx, int plus, const uint16_t *prm,
                       double *out);

}  // namespace lsg
