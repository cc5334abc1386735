/* SPDX-License-Identifier: 0BSD
 *
 * Melt-pool surrogate (no reference counterpart): a 6-layer conv3x3 network
 * over 128 x 128 windows, trained and evaluated by the kernels below.  Its
 * inputs and target in the two-level run are those of
 * include/lpbf_multirate.h (4 input channels; output: the fine 16-step
 * residual over the prolonged coarse increment, / 100 K).
 *
 *   input  x0[b][y][x][32] bf16 (NHWC): channels 0..3 the inputs, 4..31 = 0
 *   L1..L5 z = conv3x3(a_prev) + bias, a = gelu(z) (tanh form), 32 channels,
 *          bf16 weights and activations, FP32 accumulation (MFMA 16x16x32)
 *   L6     y = conv3x3(a5) + bias, FP32
 *   loss   mean over the batch's pixels of (y - target)^2
 *
 * Convolutions use zero padding at the window border.  Parameters are one
 * FP32 vector of LPBF_SG_NPARAM values: W1[9][4][32] (tap, in, out), b1[32],
 * then for L2..L5 W[9][32][32], b[32], then W6[9][32], b6.  Tap t reads the
 * input at (y + t/3 - 1, x + t%3 - 1).
 *
 * Every launch is asynchronous on `stream`; invalid arguments record
 * LPBF_ERR_INVALID_ARGUMENT (lpbf_get_last_error).
 */
#ifndef LPBF_SURROGATE_H
#define LPBF_SURROGATE_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include "lpbf_step_const.h"

#define LPBF_SG_WIN    128
#define LPBF_SG_CH     32
#define LPBF_SG_NPARAM 38465
#define LPBF_SG_STEPS  16

#ifdef __cplusplus
extern "C" {
#endif

/* Device bytes of the bf16 weight pack and of the per-batch workspace
 * (activations, and with training != 0 the backward buffers). */
size_t lpbf_sg_pack_bytes(void);
size_t lpbf_sg_work_bytes(int batch, int training);

/* FP32 parameters -> bf16 pack used by the kernels (forward and transposed
 * copies, FP32 biases). */
void lpbf_sg_pack(const float* d_params, void* d_pack, hipStream_t stream);

/* y[b][y][x] (FP32) for `batch` windows x0 (layout above). */
void lpbf_sg_forward(const void* d_pack, const void* d_x0, int batch, void* d_work,
                     float* d_y, hipStream_t stream);

/* Forward, loss against target[b][y][x] and backward: d_grad receives the
 * LPBF_SG_NPARAM gradient of the mean loss, *d_loss the loss (FP64), d_y the
 * prediction (may be NULL). */
void lpbf_sg_loss_grad(const void* d_pack, const void* d_x0, const float* d_target,
                       int batch, void* d_work, float* d_grad, double* d_loss, float* d_y,
                       hipStream_t stream);

/* Adam step t (1-based) on n parameters. */
void lpbf_sg_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int n, int t,
                  float lr, float beta1, float beta2, float eps, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LPBF_SURROGATE_H */
