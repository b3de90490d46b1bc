// fg_mlp.h -- parameter layout and launchers of the per-pixel error model
// (fg_mlp.hip).
#ifndef FG_MLP_H
#define FG_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

// One f32 array: W1 [16][64] (input-major; inputs 12..15 are padding), b1
// [64], W2 [64][64] (input-major), b2 [64], w3 [64], b3.
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + 16 * 64;
constexpr int kMlpW2 = kMlpB1 + 64;
constexpr int kMlpB2 = kMlpW2 + 64 * 64;
constexpr int kMlpW3 = kMlpB2 + 64;
constexpr int kMlpB3 = kMlpW3 + 64;
constexpr int kMlpParams = kMlpB3 + 1;

hipError_t launch_features(const float *u, int in_w, int in_h, int out_w, int out_h, float zoom, float radius,
                           float sd, float sigma, int channel, uint16_t *X, hipStream_t s);
hipError_t launch_mlp_forward(const float *P, const uint16_t *X, long long n, float *out, hipStream_t s);
// loss_kind 0: squared error to target; 1: Gaussian negative log-likelihood
// of the residual target under the predicted log std.
hipError_t launch_mlp_grad(const float *P, const uint16_t *X, const float *target, long long n, float inv_total,
                           int loss_kind, float *grad, float *loss, hipStream_t s);
hipError_t launch_budget(const float *pred, const uint16_t *X, long long n, float c, const float *dc, int nmin,
                         int nmax, uint16_t *budget, hipStream_t s);
// The budget scale search (see fg_mlp.hip): writes out[0] = c for s =
// exp(pred) * s_mult and out[1] = c * s_mult (the scale of exp(pred));
// state_buf holds budget_scale_state_bytes().
size_t budget_scale_state_bytes();
hipError_t launch_budget_scale(const float *pred, long long n, float s_mult, double target, int nmin, int nmax,
                               int n_ref, int rounds, long long max_rows, void *state_buf, float *out, hipStream_t s);
hipError_t launch_mlp_adam(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2, float eps,
                           float bc1, float bc2, hipStream_t s);

}  // namespace fg

#endif
