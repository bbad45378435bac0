/* fastiso_surrogate.h -- C ABI of the learned surrogate of libfastiso_hip.so.
 *
 * A Fourier-neural-operator network that predicts the viscous displacement increment of the
 * laterally variable ELVA model over a macro step from [u, load anomaly, log10 viscosity,
 * lithosphere thickness]. It has no counterpart in the reference (FastIsostasy.jl has no learned
 * component); it is trained only on trajectories of the plain solver (include/fastiso.h), and the
 * plain solver re-checks it (the driver is fastiso.surrogate in the Python package).
 *
 * Network (all per-cell operations are matrix products on BF16 MFMA with FP32 accumulation):
 *   v0 = x P + p                                   lift, in_channels -> width
 *   z_l = v_l W_l + b_l + S_l(v_l),  v_{l+1} = gelu(z_l) (no gelu after the last layer)
 *   y = gelu(v_L Q1 + q1) Q2 + q2                  projection, width -> proj_hidden -> 1
 * S_l is the truncated spectral convolution: the real 2D DFT of each channel restricted to
 * kx = 0..modes_x-1 and ky = 0..modes_y-1, ny-modes_y..ny-1, a complex width x width matrix per
 * retained mode, and the inverse transform back to the grid, normalised by 1/(nx ny) so that
 * the weights do not depend on the grid size.
 *
 * Parameter vector (FP32), in this order: P[in][width], p[width]; per layer W[width][width],
 * b[width], R[modes_x][2 modes_y][width(in)][width(out)][re, im]; Q1[width][hidden], q1[hidden],
 * Q2[hidden], q2[1].
 * Grids: cell (i, j) of an nx x ny grid at i + j nx; inputs are cell-major with the channel
 * fastest (x[(i + j nx) in_channels + c]).
 */
#ifndef FASTISO_SURROGATE_H
#define FASTISO_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fi_fno_config {
    int32_t in_channels;   /* 4 */
    int32_t width;         /* 32 */
    int32_t modes_x;       /* 16 */
    int32_t modes_y;       /* 16 (2 modes_y rows of ky are kept) */
    int32_t layers;        /* 4 */
    int32_t proj_hidden;   /* 64 */
    int32_t device;
    int32_t reserved;
    uint64_t seed;         /* parameter initialisation */
} fi_fno_config;

typedef struct fi_surrogate fi_surrogate;

void fi_fno_default_config(fi_fno_config* cfg);
int fi_surrogate_create(const fi_fno_config* cfg, fi_surrogate** out);
void fi_surrogate_destroy(fi_surrogate* s);
const char* fi_surrogate_last_error(const fi_surrogate* s);
int64_t fi_surrogate_num_params(const fi_surrogate* s);
int fi_surrogate_get_params(fi_surrogate* s, float* host, int64_t n);
int fi_surrogate_set_params(fi_surrogate* s, const float* host, int64_t n);

/* Grid the next calls work on (builds the DFT matrices and activation buffers once per size). */
int fi_surrogate_set_grid(fi_surrogate* s, int32_t nx, int32_t ny);

/* y = network(x) on the current grid; host arrays. */
int fi_surrogate_forward(fi_surrogate* s, const float* x, float* y);

/* Loss 1/(nx ny) sum (y - target)^2 of one sample and its gradient with respect to every
 * parameter (host arrays; grad has fi_surrogate_num_params entries). */
int fi_surrogate_loss_grad(fi_surrogate* s, const float* x, const float* target, double* loss, float* grad);

/* The same for n samples side by side (x: n * cells * in_channels, target: n * cells): loss and
 * gradient summed over the samples, which go through the network together (up to 16 per pass). */
int fi_surrogate_loss_grad_n(fi_surrogate* s, int32_t n, const float* x, const float* target, double* loss,
                             float* grad);

/* Training set on the device: nsamples samples of the current grid (x: nsamples * cells *
 * in_channels, y: nsamples * cells). Replaces any previous set. */
int fi_surrogate_set_data(fi_surrogate* s, int32_t nsamples, const float* x, const float* y);

/* steps Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, decoupled weight decay wd) on mini-batches
 * of batch samples drawn with a seeded generator; learning rate lr (cosine decay to lr_min over
 * the steps). loss_out (optional, steps entries) receives each step's mean mini-batch loss. */
int fi_surrogate_train(fi_surrogate* s, int32_t steps, int32_t batch, float lr, float lr_min, float wd,
                       uint64_t seed, double* loss_out);

/* Device time of all kernels launched by the surrogate so far (HIP events), in ms. */
double fi_surrogate_device_ms(const fi_surrogate* s);

#ifdef __cplusplus
}
#endif
#endif /* FASTISO_SURROGATE_H */
