/*
 * lbm_surrogate.h -- C ABI of the warm-start surrogate: a convolutional
 * encoder-decoder that maps (fluid mask, x ramp, tau, density drop) on an
 * H x W grid to a normalised steady (density, momentum) field, trained and
 * evaluated on the GPU by this library's own HIP kernels.
 *
 * No counterpart in the reference (its solver always starts from rest,
 * include/lbm/core/domain_initialization.hpp:311-403); callers use the
 * prediction only to warm-start a lattice from its equilibrium, and keep the
 * plain solve as the arbiter of convergence and as the fallback.
 *
 * Network (NCHW float32; every conv 3x3, zero padding, ReLU except the last;
 * pool = 2x2 mean; upadd = nearest 2x upsampling plus skip):
 *   enc: 4->16->16 | pool 16->32->32 | pool 32->48->48 | pool 48->64->64 | pool 64->64
 *   dec: upadd + 64->48 | upadd + 48->32 | upadd + 32->16 | upadd + 16->16, 16->3
 * 14 convolutions, 201 955 parameters: per conv, weights [Cout][Cin][3][3]
 * then bias [Cout].  H and W must be positive multiples of 16.
 * Loss: sum over fluid nodes (input channel 0 = 1) of the squared error of
 * the 3 outputs, divided by 3 x the fluid node count.  Optimiser: Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).
 *
 * Every function returns LBM_OK (0) or an lbm_hip.h error code; the message
 * is in lbm_last_error().  Calls on one handle must not overlap.
 */
#ifndef LBM_SURROGATE_H
#define LBM_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_surrogate lbm_surrogate;

enum { LBM_SURROGATE_INPUTS = 4, LBM_SURROGATE_OUTPUTS = 3, LBM_SURROGATE_CONVS = 14 };

size_t lbm_surrogate_param_count(void);

/* Buffers for batches of up to max_batch grids of height x width. */
int lbm_surrogate_create(uint32_t height, uint32_t width, uint32_t max_batch, lbm_surrogate **out);
void lbm_surrogate_destroy(lbm_surrogate *s);

/* He-normal weights from a splitmix64 stream of `seed`, zero biases; resets Adam. */
int lbm_surrogate_init_params(lbm_surrogate *s, uint64_t seed);
/* Host copies of the flat parameter vector; set resets Adam. */
int lbm_surrogate_set_params(lbm_surrogate *s, const float *params, size_t count);
int lbm_surrogate_get_params(lbm_surrogate *s, float *params, size_t count);

/* Forward pass: inputs (batch x 4 x H x W) -> outputs (batch x 3 x H x W),
 * host arrays; any batch size (processed in chunks of max_batch). */
int lbm_surrogate_predict(lbm_surrogate *s, const float *inputs, uint32_t batch, float *outputs);
/* The same with the convolutions on the matrix cores in BF16 (inputs of each
 * convolution and its weights rounded to BF16, FP32 accumulation; pooling,
 * upsampling and the skips in FP32). */
int lbm_surrogate_predict_bf16(lbm_surrogate *s, const float *inputs, uint32_t batch, float *outputs);

/* Loss and the gradient of every parameter for one host batch (batch <= max_batch). */
int lbm_surrogate_loss_grad(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t batch,
                            double *loss, float *grad, size_t count);

/* One Adam step on one host batch; *loss is the loss before the step. */
int lbm_surrogate_train_step(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t batch,
                             float learning_rate, double *loss);

/* Device-resident training set: room for `capacity` samples, then append. */
int lbm_surrogate_dataset_reserve(lbm_surrogate *s, size_t capacity);
int lbm_surrogate_dataset_add(lbm_surrogate *s, const float *inputs, const float *targets, uint32_t count);
size_t lbm_surrogate_dataset_size(const lbm_surrogate *s);

/* `steps` Adam steps on minibatches of `batch` samples drawn without
 * replacement (reshuffled each epoch by a splitmix64 stream of `seed`, taken
 * on the first call after the parameters or the dataset were set).  The
 * learning rate follows one schedule of `schedule_steps` steps across
 * successive calls: linear warm-up over the first min(200, schedule_steps/10)
 * steps, then cosine from lr_max down to lr_min.  loss_history (steps
 * values, or NULL) receives each step's loss. */
int lbm_surrogate_train(lbm_surrogate *s, uint32_t steps, uint32_t schedule_steps, uint32_t batch, float lr_max,
                        float lr_min, uint64_t seed, float *loss_history);

/* Milliseconds of the last predict / train call (HIP events). */
int lbm_surrogate_last_ms(const lbm_surrogate *s, double *ms);

#ifdef __cplusplus
}
#endif

#endif
