/* SPDX-License-Identifier: 0BSD
 *
 * Melt-pool surrogate (no reference counterpart): a 6-layer conv3x3 network
 * that advances a 128 x 128 window centred on the laser by 16 steps.
 *
 *   input  x0[b][y][x][32] bf16 (NHWC): channel 0 = (T_t - T_ref)/3000,
 *          1 = laser source map (mean over the 16 steps of exp(-2 r^2/r0^2),
 *          beam-off steps count 0), 2 = P/400, 3 = V/2, 4..31 = 0
 *   L1..L5 z = conv3x3(a_prev) + bias, a = gelu(z) (tanh form), 32 channels,
 *          bf16 weights and activations, FP32 accumulation (MFMA 16x16x32)
 *   L6     y = conv3x3(a5) + bias, FP32: the predicted (T_{t+16} - T_t)/3000
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

/* One window of an FP32 field (theta = T - T_ref, pitch ld): origin (ox, oy)
 * such that window cell (i, j) is field cell (ox + j, oy + i), with
 * 0 <= ox <= nx - 128 and 0 <= oy <= ny - 128; the laser of steps
 * step .. step+15 from `plan`; dx, dy, 1/r0^2 of the run; P, V. */
typedef struct lpbf_sg_window {
    const float*     theta;
    float*           theta_out;   /* lpbf_sg_apply: the field to write */
    const LaserTick* plan;
    double           T_ref, dx, dy, inv_r0_sq, P, V;
    int              ld, nx, ny, ox, oy, step;
} lpbf_sg_window;

/* x0 (bf16, layout above) and T0[b][y][x] = theta at the window (FP32) for
 * n windows described by d_win (device array). */
void lpbf_sg_gather(const lpbf_sg_window* d_win, int n, void* d_x0, float* d_T0,
                    hipStream_t stream);

/* target[b][y][x] = (theta_now - T0)/3000 over each window (theta read from
 * d_win[b].theta, typically the field 16 steps after lpbf_sg_gather). */
void lpbf_sg_target(const lpbf_sg_window* d_win, int n, const float* d_T0, float* d_target,
                    hipStream_t stream);

/* Writes the prediction into each window of d_win[b].theta_out, blended
 * with the field's current values over a border of `overlap` cells (weight
 * of the prediction rising linearly from 0 at the window edge to 1 at
 * `overlap` cells in), unless d_skip[b] != 0 (d_skip may be NULL).
 * d_maxdiff[b] (may be NULL; zeroed by the caller) receives the largest
 * |prediction - current field| in kelvin over the window before blending. */
void lpbf_sg_apply(const lpbf_sg_window* d_win, int n, const float* d_T0, const float* d_y,
                   int overlap, const int* d_skip, float* d_maxdiff, hipStream_t stream);

#ifdef __cplusplus
}
static_assert(sizeof(lpbf_sg_window) == 96, "lpbf_sg_window layout");
#else
_Static_assert(sizeof(lpbf_sg_window) == 96, "lpbf_sg_window layout");
#endif

#endif /* LPBF_SURROGATE_H */
