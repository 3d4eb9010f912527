/*
 * cgs_unet.h -- C ABI of the learned CG warm start in libcgsense_hip.so.
 *
 * No counterpart in codeare: codeare's CGLS (src/optimisation/CGLS.hpp:30-59)
 * always starts from zero.  These entry points build, train and run a 2D U-Net
 * (3x3 convolutions + ReLU, 2x2 max-pool down, nearest 2x up + skip
 * concatenation, a 1x1 head with a residual connection) whose output is passed
 * as x0 to cgs_solve.  Convolutions run on BF16 MFMA with FP32 accumulation;
 * parameters are FP32 masters updated by Adam.
 *
 * Images are complex64 (batch, H, W) with W fastest -- a codeare column-major
 * (N0, N1) image is W = N0, H = N1.  By default pointers are host memory; with
 * device_io != 0 they are device memory.  Every call returns when its work is
 * done.  A network object is not re-entrant.
 */
#ifndef CGS_UNET_H
#define CGS_UNET_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgs_unet_s* cgs_unet;

/* levels: 2..6 resolution levels (4 = the 256^2 warm-start network);
 * base_channels: 32 (channels double per level); H, W multiples of
 * 16 * 2^(levels-1); max_batch bounds every call's batch.  He-normal weights
 * from `seed`, zero biases, zero head (the untrained network is the identity). */
int cgs_unet_create(cgs_unet* net, int H, int W, int levels, int base_channels, int max_batch,
                    unsigned long long seed);
int cgs_unet_destroy(cgs_unet net);

/* parameter vector (FP32): per conv layer W[tap][co][ci] (tap = 3*ky + kx)
 * then bias[co], layers in order enc(level 0 a, b, ..., level L-1 a, b),
 * dec(level L-2 a, b, ..., level 0 a, b), then head W[2][32], head bias[2] */
int cgs_unet_num_params(cgs_unet net, long long* count);
int cgs_unet_get_params(cgs_unet net, float* host);
int cgs_unet_set_params(cgs_unet net, const float* host);

/* y = net(x) */
int cgs_unet_forward(cgs_unet net, const void* x, void* y, int batch, int device_io);
/* loss = mean over real components of (net(x) - target)^2; gradients of the
 * loss w.r.t. every parameter are left in the net (read with cgs_unet_get_grads) */
int cgs_unet_gradients(cgs_unet net, const void* x, const void* target, int batch, int device_io,
                       double* loss);
int cgs_unet_get_grads(cgs_unet net, float* host);
/* cgs_unet_gradients + one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) */
int cgs_unet_train_step(cgs_unet net, const void* x, const void* target, int batch, double lr,
                        int device_io, double* loss);
int cgs_unet_reset_optimizer(cgs_unet net);

/* kernel timing: enable = 1 starts timing every launch with HIP events (and
 * clears the totals), 0 stops, -1 leaves it as is.  stats8 (may be null)
 * receives the totals so far: conv forward ms, conv forward flop, conv
 * input-gradient ms, flop, conv weight-gradient ms, flop, other kernels ms, 0. */
int cgs_unet_profile(cgs_unet net, int enable, double* stats8);

#ifdef __cplusplus
}
#endif

#endif /* CGS_UNET_H */
