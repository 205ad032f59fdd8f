// Cost-model MLP shared declarations (mlp.hip, learned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trust {

struct Ctx;

constexpr int kMlpHidden = 32;
template <int H>
constexpr int kMlpParams = H * H + 6 * H + 1;

hipError_t mlp_forward(const float *params, const float *x, long long n, float *y, hipStream_t s);
hipError_t mlp_loss_grad(const float *params, const float *x, const float *t, long long n,
                         float *partial, float *grad_out, hipStream_t s);
hipError_t adam_step(float *params, const float *grad, float *m, float *v, int P, float lr, int t,
                     hipStream_t s);
hipError_t launch_gather_samples(const float *x, const float *t, const long long *idx, long long n,
                                 float *bx, float *bt, hipStream_t s);
void mlp_init_host(uint64_t seed, float *params);
int mlp_train_device(Ctx *ctx, float *d_params, const float *d_x, const float *d_t, long long n,
                     int epochs, int batch, float lr, uint64_t seed, double *loss_out);

// Feature transform shared by training and inference: log2(1 + raw) / 32.
constexpr float kFeatScale = 1.0f / 32.0f;

}  // namespace trust
