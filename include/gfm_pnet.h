/*
 * gfm_pnet.h -- C-ABI of the learned initial guess for the pressure-
 * correction (p') solve.
 *
 * The reference starts every p' solve from zero (FZ = 0,
 * C/Gslv3v.f90:157-161) and relaxes a fixed 5 sweeps.  The solver here can
 * instead start from the prediction of a small 3-D convolutional network
 * (a U-Net) that maps the mass source of the p' equation to p', then relax
 * to a residual tolerance, falling back to the zero guess when the
 * prediction does not lower the residual (gfm_hip.h: pcorr_mode,
 * pcorr_guess).  This header has no counterpart in the reference: it
 * trains and evaluates that network on the GPU.
 *
 * Network (fixed topology, channel widths C0..C3 chosen at creation):
 *   level 0: conv(2->C0), conv(C0->C0)           -> s0
 *   level 1: avgpool2, conv(C0->C1), conv(C1->C1) -> s1
 *   level 2: avgpool2, conv(C1->C2), conv(C2->C2) -> s2
 *   level 3: avgpool2, conv(C2->C3)               -> s3
 *   decoder: d2 = conv([up(s3), s2] -> C2), d1 = conv([up(d2), s1] -> C1),
 *            d0 = conv([up(d1), s0] -> C0)
 *   output:  y = w . d0 + b (1x1x1, linear)
 * Every conv is 3x3x3 with zero padding, bias and ReLU; up() is nearest-
 * neighbour upsampling by 2.  Activations are BF16 (round to nearest even)
 * with FP32 accumulation on the matrix cores; parameters, gradients and
 * Adam moments are FP32.  Grid dimensions must be multiples of 8.
 *
 * Input x: [batch][d][h][w][2] float, channel 0 the normalised source,
 * channel 1 the open-cell mask (1 open, 0 blocked).  Target t: [batch][d]
 * [h][w] float.  Loss: mean of (y - t)^2 over open cells of the batch.
 *
 * Flat parameter vector (float), conv layers 0..9 in the order above, each
 * W[27][Co][Cin_pad] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1), input channel
 * fastest; layer 0 pads its 2 input channels to 32 with weights that stay
 * zero) followed by bias[Co]; then the output layer w[C0], b.
 *
 * Error convention as in gfm_hip.h (0 on success, GFM_E_* otherwise,
 * message from gfm_hip_last_error()).
 */
#ifndef GFM_PNET_H
#define GFM_PNET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfm_pnet gfm_pnet;

/* Create a network on `device` with channel widths ch[0..3] (multiples of
 * 32) and buffers for up to max_batch grids of max_cells cells; weights
 * He-normal from `seed`, biases zero. */
int gfm_pnet_create(int device, const int ch[4], int max_batch, long long max_cells, uint64_t seed, gfm_pnet **out);
int gfm_pnet_destroy(gfm_pnet *net);

/* Length of the flat parameter vector and the number of parameters that
 * are not padding. */
int gfm_pnet_param_count(const gfm_pnet *net, size_t *n_total, size_t *n_effective);
int gfm_pnet_get_params(gfm_pnet *net, float *out, size_t n);
int gfm_pnet_set_params(gfm_pnet *net, const float *in, size_t n);

/* Forward pass on host arrays. */
int gfm_pnet_forward(gfm_pnet *net, int batch, int d, int h, int w, const float *x, float *y);

/* Loss and parameter gradients on host arrays; grads (nullable) receives
 * the flat gradient vector.  The gradients stay on the device for
 * gfm_pnet_adam_step. */
int gfm_pnet_loss_grad(gfm_pnet *net, int batch, int d, int h, int w, const float *x, const float *t, double *loss,
                       float *grads);

/* One Adam step (bias-corrected) with the gradients of the last
 * gfm_pnet_loss_grad or training batch. */
int gfm_pnet_adam_step(gfm_pnet *net, float lr, float beta1, float beta2, float eps);

/* Train on nsamples (x, t) pairs of one grid size: the set is copied to
 * the device once, each epoch visits it in a seeded random order in
 * batches, one Adam step (beta 0.9/0.999, eps 1e-8) per batch.
 * epoch_loss[e] receives the mean batch loss of epoch e. */
int gfm_pnet_train(gfm_pnet *net, int nsamples, int d, int h, int w, const float *x, const float *t, int epochs,
                   int batch, float lr, uint64_t seed, double *epoch_loss);

#ifdef __cplusplus
}
#endif
#endif /* GFM_PNET_H */
