// The barostat's learned screen: launchers of the BF16-MFMA MLP kernels.
#pragma once

#include "kernels.h"

namespace rmd {

constexpr int SURROGATE_INPUTS = 16;
constexpr int SURROGATE_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1;  // 5313

// Raises the dynamic-LDS limit of the three kernels (once per process).
void surrogateConfigure();
// y[k] = f(x[k]) for k < n (x: n x 16 floats).
void launchSurrogateForward(const float *params, const float *x, int n, float *y, hipStream_t s);
// Mean-squared-error loss over the n samples and its gradient (SURROGATE_PARAMS floats).
void launchSurrogateGradient(const float *params, const float *x, const float *t, int n, float *loss, float *grad,
                             hipStream_t s);
// `steps` Adam steps on mini-batches of 64 drawn with a counter hash.
void launchSurrogateTrain(float *params, float *m, float *v, const float *x, const float *t, int n, int steps, float lr,
                          unsigned int seed, int stepOffset, float *loss, hipStream_t s);

}  // namespace rmd
