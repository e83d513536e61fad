/*
 * erm_unet.h - C ABI of the learned LSRTM preconditioner in liberm.so.
 *
 * No reference counterpart: the reference has no learned component.  A
 * 4-level U-Net (channels 16/32/64/64 by default; 3x3 convolutions, two per
 * level, ReLU, 2x2 average pooling, nearest 2x upsampling, skip
 * concatenation, 3x3 output convolution plus the input as a residual) maps a
 * 2-channel image (the migrated LSRTM gradient L^T d in dvp/vp, dvs/vs) to
 * the converged LSRTM model of the same area.  Convolutions run as
 * implicit GEMMs on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
 * accumulation; master weights and Adam state are FP32, the MFMA operands
 * are the BF16 roundings (round to nearest even) of weights and activations.
 *
 * Tensors are NCHW float32 host arrays.  H and W must be multiples of 8.
 */
#ifndef ERM_UNET_H
#define ERM_UNET_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct erm_unet erm_unet;

const char* erm_unet_last_error(void);
/* c1..c4: channels per level (16, 32, 64, 64); seed for He initialisation. */
int   erm_unet_create(int device, int c1, int c2, int c3, int c4, unsigned long long seed, erm_unet** out);
int   erm_unet_destroy(erm_unet* u);
/* Number of float32 parameters (weights then biases, layer by layer). */
size_t erm_unet_num_params(const erm_unet* u);
int   erm_unet_get_params(erm_unet* u, float* dst);
int   erm_unet_set_params(erm_unet* u, const float* src);
/* y = net(x), x and y [n][2][h][w]. */
int   erm_unet_forward(erm_unet* u, const float* x, int n, int h, int w, float* y);
/* One loss evaluation and backward pass: loss = mean((net(x) - t)^2);
 * grad receives d loss / d params (same order as erm_unet_get_params). */
int   erm_unet_loss_grad(erm_unet* u, const float* x, const float* t, int n, int h, int w, float* loss, float* grad);
/* One Adam step (Kingma & Ba 2015; beta1 0.9, beta2 0.999, eps 1e-8) on a batch. */
int   erm_unet_train_step(erm_unet* u, const float* x, const float* t, int n, int h, int w, float lr, float* loss);
/* Training over a set of patches for `epochs` epochs with batches of `batch`
 * (shuffled with `seed`); losses[epochs] receives the mean loss per epoch. */
int   erm_unet_train(erm_unet* u, const float* x, const float* t, int count, int h, int w, int batch, int epochs,
                     float lr, unsigned long long seed, float* losses);
/* Preconditioner adaptor for erm2d_lsrtm_cg (erm2d_precond_fn, user = the
 * net): bind the model image shape [3][nx][nz] (z fastest) once, then
 * z = s * net(g / s) on the (dvp/vp, dvs/vs) planes, with s the RMS of g over
 * those planes and the image zero-padded to multiples of 8 (h = x, w = z);
 * the drho/rho plane passes through (times w when weights are bound, see
 * erm_unet_bind_weights).  Degree-1 homogeneous, so CG's line
 * search sees the same directions for any scaling of g. */
int   erm_unet_bind_model(erm_unet* u, int nx, int nz);
int   erm_unet_precond(const float* g, float* z, void* user);
/* Optional per-cell input weights w [nx][nz] for erm_unet_precond (after
 * erm_unet_bind_model, which clears them; w = NULL clears): the net is
 * applied to w * g instead of g, so a diagonal preconditioner (e.g.
 * erm2d_illumination-based) and the learned correction compose into one
 * approximate inverse Hessian z = net(w * g).  Still degree-1 homogeneous. */
int   erm_unet_bind_weights(erm_unet* u, const float* w);
/* Milliseconds of device time of the last forward / train call and the
 * number of MFMA floating-point operations it issued (2*M*N*K per GEMM,
 * padded shapes), for utilisation reporting. */
int   erm_unet_last_stats(const erm_unet* u, double* ms, double* mfma_flops);

#ifdef __cplusplus
}
#endif
#endif
