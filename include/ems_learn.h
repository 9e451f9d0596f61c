/*
 * ems_learn.h -- the learned early stop: a causal dilated 1-D convolution
 * network that extrapolates port voltage/current traces, trained and run
 * on the device by hand-written kernels.
 *
 * No counterpart in the reference: its run loop (src/engine/run.rs:424-540,
 * drive) always steps until the end criterion. Callers that want an
 * early-stopped sweep run each model for a fraction of the steps, feed the
 * observed traces to the network, and fall back to the full run when the
 * prediction does not decay to the end criterion within its horizon.
 *
 * Network: `layers` layers of causal convolution with 9 taps and dilation
 * dilation[l]; layer 0 maps cin channels to `channels` (tanh), the middle
 * layers are residual (h += tanh(conv(h))), the last maps to cout channels
 * without activation. Weights are kept as f32 master parameters (the
 * optimizer's) and a bf16 copy (round to nearest even) that every forward
 * and backward pass reads; arithmetic is f32.
 *
 * Parameter vector (f32): for each layer l, the weights [out][in][tap]
 * then the biases [out].
 * Sequences: [sample][channel][time] f32, time fastest.
 * Loss: mean over samples, output channels and t >= t0 of (out - y)^2,
 * plus an optional quadratic (spectral) term, ems_cn_set_spectral.
 * Optimizer: Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected),
 * optionally on a norm-clipped gradient.
 *
 * Return codes as ems_gpu.h (EMS_OK, EMS_E_INVALID, EMS_E_DEVICE).
 */
#ifndef EMS_LEARN_H
#define EMS_LEARN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ems_cn ems_cn;

/* A network with seeded initial weights (normal, std 1/sqrt(fan_in), the
 * output layer 0.1 of that; biases 0). */
int ems_cn_create(int cin, int cout, int channels, int layers, const int32_t *dilation,
                  uint64_t seed, ems_cn **out);
void ems_cn_destroy(ems_cn *n);
const char *ems_cn_last_error(const ems_cn *n);
uint64_t ems_cn_param_count(const ems_cn *n);

/* Replaces the parameters (and resets the optimizer state). */
int ems_cn_set_params(ems_cn *n, const float *params, uint64_t count);
int ems_cn_get_params(ems_cn *n, float *params, uint64_t count);

/* Adds a quadratic term to the loss: for every sample and output channel,
 * weight * d^T K d / (samples * channels), with d the error (output -
 * target) over samples [t0, t0 + len) and K a symmetric len x len matrix
 * (row-major, copied). A K built from cosines at chosen frequencies makes
 * it the error of the spectrum there. weight 0 removes the term. */
int ems_cn_set_spectral(ems_cn *n, const float *K, int t0, int len, float weight);

/* Uploads a training set: x [count][cin][T], y [count][cout][T]. */
int ems_cn_set_data(ems_cn *n, const float *x, const float *y, int count, int T);

/* One optimizer step on the samples idx[0..B) of the training set, loss
 * over t >= t0; the gradient is scaled to 2-norm `clip` when it is longer
 * (clip <= 0: no clipping). *loss and *grad_norm (either may be NULL)
 * receive the batch loss before the step and the gradient's norm before
 * clipping. */
int ems_cn_train_batch(ems_cn *n, const int32_t *idx, int B, int t0, float lr, float clip,
                       float *loss, float *grad_norm);

/* out [B][cout][T] = network(x [B][cin][T]). */
int ems_cn_forward(ems_cn *n, const float *x, int B, int T, float *out);

/* Loss and its gradient with respect to the parameters (grad: NULL or
 * param_count floats) on the batch (x, y), without an optimizer step. */
int ems_cn_loss_grad(ems_cn *n, const float *x, const float *y, int B, int T, int t0,
                     float *loss, float *grad);

#ifdef __cplusplus
}
#endif

#endif
