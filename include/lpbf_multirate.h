/* SPDX-License-Identifier: 0BSD
 *
 * Two-level (multirate) run with a learned melt-pool window (no reference
 * counterpart; the reference advances every cell of the plate with the FD
 * step, include/step_task.hpp:54-206).
 *
 * The far field lives on a coarse grid: 2x2 fine cells per coarse cell
 * (nx, ny even), four fine time steps per coarse step.  Its explicit step
 * has the fine step's stability number (alpha * dt / dx^2 is unchanged), so
 * it costs 1/16 of the fine plate per simulated second.  The melt pool lives
 * in a fine window of 128 x 128 cells that follows the laser (96 cells of
 * its trail, 32 ahead of it), advanced 16 fine steps at a time:
 *
 *   base      the window at t plus the coarse grid's 16-step increment
 *             (coarse cell values at t hold the window's restriction),
 *             prolonged to the fine cells (cell-centred bilinear);
 *   network   the surrogate (include/lpbf_surrogate.h) predicts the fine
 *             residual (fine increment - prolonged coarse increment) / 100 K
 *             from four input channels (below); its weight rises from 0 at
 *             the window edge to 1 at `overlap` cells in;
 *   FD window the fine step itself on the window plus 16 cells of coarse
 *             field on every side, prolonged conservatively (per axis the
 *             quadratic through three coarse cell means; the 16 steps cannot
 *             carry the margin's edge error into the window), stepped in LDS by one
 *             launch: the network's judge every `recheck_every` calls and
 *             its fallback;
 *   coarse    the plate's coarse step takes lpbf_mr_coarse_coeffs (four fine
 *             steps' loss and diffusion in one step); the window-local coarse
 *             step that feeds the network keeps the plain coarse step (dt x 4
 *             in one explicit step), the increment it was trained on;
 *   coupling  after each window advance its cells overwrite their coarse
 *             cells (2x2 mean, all but `overlap`/2 coarse cells at the
 *             window edge), and when the window moves the cells it gains are
 *             prolonged (quadratic) from the coarse grid.  The network's
 *             channels and base use the bilinear form it was trained with.
 *
 * Network input channels (bf16, the 4 of the 32-channel layer-1 input):
 *   0  (theta - P R theta) / 100 K: the fine-scale part of the window field
 *      (R = 2x2 mean over the window, P = cell-centred bilinear prolongation
 *      within the window);
 *   1  theta / 3000 K;
 *   2  prolonged coarse 16-step increment / 1000 K;
 *   3  laser source map: mean over the 16 steps of q * exp(-2 r^2 / r0^2).
 *
 * Every launch is asynchronous on the given stream; return codes are 0 or a
 * hipError_t / LPBF_ERR_INVALID_ARGUMENT, also recorded for
 * lpbf_get_last_error().
 */
#ifndef LPBF_MULTIRATE_H
#define LPBF_MULTIRATE_H

#include <hip/hip_runtime_api.h>
#include "lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One 128 x 128 window for lpbf_mr_gather / lpbf_mr_target.  Window cell
 * (i, j) is fine plate cell (ox + j, oy + i) and is read from
 * fine[(fy + i) * ldf + fx + j] (the plate itself, fx = ox, fy = oy, or a
 * window buffer, fx = fy = 0).  The coarse fields c0 (time t) and c1
 * (t + 16 fine steps) are whole coarse plates of nxc x nyc cells, pitch
 * ldc, or c1 (ld1 > 0) the block of coarse cells around the window that
 * the prolongation reads. */
typedef struct lpbf_mr_window {
    const float*     fine;
    const float*     c0;
    const float*     c1;
    const LaserTick* plan;     /* fine laser ticks; steps step .. step+15 */
    double           dx, dy, inv_r0_sq;
    int              ldf, fx, fy;
    int              ox, oy;   /* even */
    int              ldc, nxc, nyc;
    int              step;
    int              ld1, c1x, c1y;  /* ld1 > 0: c1 holds only coarse cells from (c1x, c1y),
                                        cell (u, v) at c1[(v - c1y) * ld1 + u - c1x] */
    int              pad_;
} lpbf_mr_window;

/* coarse[v * ldc + u] = mean of the fine 2x2 block (2u..2u+1, 2v..2v+1);
 * nx and ny even. */
void lpbf_mr_restrict(const float* fine, int ldf, int nx, int ny, float* coarse, int ldc,
                      hipStream_t stream);

/* For n windows (`fine` at time t): network input x0 (bf16
 * [n][128][128][32], channels above, 4..31 zero), D[n][128][128] = prolonged
 * coarse increment (K) and T0[n][128][128] = theta at t (D and T0 may be
 * NULL). */
void lpbf_mr_gather(const lpbf_mr_window* d_win, int n, void* d_x0, float* d_D, float* d_T0,
                    hipStream_t stream);

/* Training target for n windows whose `fine` now holds the field at t + 16:
 * target[n][128][128] = (theta(t+16) - T0 - D) / 100 K. */
void lpbf_mr_target(const lpbf_mr_window* d_win, int n, const float* d_T0, const float* d_D,
                    float* d_target, hipStream_t stream);

/* Window origins of a run (fine cells, even, inside the plate): for each
 * block b = 0..n_blocks (fine step 16 b) the laser of h_plan[16 b] sits at
 * 3/4 of the window along its direction of travel in x (from h_plan[16 b]
 * to h_plan[16 (b + 1)]; the last direction seen while x does not move, +x
 * before any motion), centred across it, so the window holds 96 cells of
 * the trail and 32 ahead.  h_plan holds 16 * n_blocks + 1 ticks. */
int  lpbf_mr_window_origins(const LaserTick* h_plan, int n_blocks, int nx, int ny, double dx,
                            double dy, int* ox, int* oy);

/* The coarse plate's step coefficients for the fine constants k (host
 * arithmetic, no device work).  One coarse step (dt x 4, dx x 2) reproduces
 * four fine explicit steps: with the fine step theta' = Phi(theta) + D theta
 * (Phi(theta) = theta - dt loss(theta), D the diffusion), the centre part is
 * Phi composed four times (quartic, FP64) less 2 (cx + cy) theta, and the
 * diffusion numbers are the fine ones times (1 - dt loss'(0))^3.
 * out[0] = cx, out[1] = cy, out[2 + i] = coefficient of theta^i of the centre
 * polynomial (theta' = P(theta) + cy (up + down) + cx (left + right) + source). */
int  lpbf_mr_coarse_coeffs(const StepConst* k, double T_ref, double* out7);

/* ---- the two-level run ----------------------------------------------------- */

typedef struct lpbf_mr_config {
    StepConst k;            /* fine plate (baseline physics, mirror edges); x_l.. unused */
    double    T_ref;        /* theta origin [K] (the initial field is uniform T_ref) */
    double    tol_K;        /* recheck tolerance (max |network - FD| over the window) */
    int       n_steps;      /* fine steps, a multiple of 16; the plan holds n_steps + 16 */
    int       recheck_every;/* network calls per recheck (>= 1) */
    int       overlap;      /* seam cells (even, 2..32) */
    int       use_network;  /* 0: the FD window advances every block (no network) */
    int       coarse_depth; /* coarse temporal depth 1..4; 0: depth 4 (one launch per
                               block) with the band height timed at creation */
    int       coarse_rows;  /* coarse band height (0 = built-in) */
} lpbf_mr_config;

typedef struct lpbf_mr_counters {
    long long blocks;       /* 16-step blocks advanced */
    long long calls;        /* network calls */
    long long applied;      /* network results kept */
    long long rechecks;
    long long fallbacks;    /* rechecks above tol_K (their FD result kept) */
    long long fd_blocks;    /* blocks whose window the FD patch advanced */
    double    max_recheck_K;/* largest recheck difference seen */
    double    max_applied_recheck_K; /* largest recheck difference of a kept result */
} lpbf_mr_counters;

typedef struct lpbf_mr_run lpbf_mr_run;

/* h_plan: n_steps + 16 fine ticks (host).  d_pack: the surrogate's weight
 * pack (lpbf_sg_pack; device, kept by the caller for the run's lifetime),
 * NULL when use_network == 0.  Allocates every buffer (synchronises
 * `stream`). */
int  lpbf_mr_create(const lpbf_mr_config* cfg, const LaserTick* h_plan, const void* d_pack,
                    hipStream_t stream, lpbf_mr_run** out);
void lpbf_mr_destroy(lpbf_mr_run* r);
/* Advances n_blocks blocks of 16 fine steps (fewer at the end of the plan);
 * synchronises `stream` at each recheck to take the fallback decision. */
int  lpbf_mr_advance(lpbf_mr_run* r, int n_blocks);
/* The run's field on the fine grid (FP64 T, nx*ny row-major, device): the
 * window's cells where it lies, the coarse field's quadratic prolongation
 * elsewhere. */
int  lpbf_mr_field(const lpbf_mr_run* r, double* d_T);
int  lpbf_mr_get_counters(const lpbf_mr_run* r, lpbf_mr_counters* out);
/* Up to max recheck differences in kelvin, in order; returns their count. */
int  lpbf_mr_recheck_diffs(const lpbf_mr_run* r, float* out, int max);

#ifdef __cplusplus
}
static_assert(sizeof(lpbf_mr_window) == 112, "lpbf_mr_window layout");
static_assert(sizeof(lpbf_mr_config) == 256, "lpbf_mr_config layout");
#else
_Static_assert(sizeof(lpbf_mr_window) == 112, "lpbf_mr_window layout");
_Static_assert(sizeof(lpbf_mr_config) == 256, "lpbf_mr_config layout");
#endif

#endif /* LPBF_MULTIRATE_H */
