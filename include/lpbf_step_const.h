/* SPDX-License-Identifier: 0BSD
 *
 * Plain-old-data records that cross the C ABI of the MI355X step library.
 *
 * StepConst and Partial keep the field order, types and byte layout of the
 * reference records (include/lpbf_step_const.hpp:19-34 in hedgehog_lpbf), so a
 * caller that fills a StepConst with the reference's make_const() can hand it
 * to this library unchanged.  The remaining records are new: they carry the
 * per-step laser plan, the fused per-step statistics (peak temperature,
 * melt-pool extent and area, energy balance) and the batched-instance
 * descriptors of the FP32 path.
 *
 * This header is C99 and C++ compatible and names no HIP or CUDA type.
 */
#ifndef LPBF_STEP_CONST_H
#define LPBF_STEP_CONST_H

#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Every scalar one explicit-Euler update needs (reference
 * include/lpbf_step_const.hpp:21-32).  216 bytes. */
typedef struct StepConst {
    int    nx, ny;
    double dx, dy, inv_dx2, inv_dy2, inv_r0_sq;
    double alpha, K, rho, cp, rho_cp, inv_h, h, dt;
    double peak_S, eta, P;
    double hConv, Tamb, Tsub, dSub, eps;
    bool   radiation;
    double kmult;
    double x_l, y_l;   /* laser position used by this step [m] */
    double q_scale;    /* source multiplier, 0 while the beam is off */
} StepConst;

/* Block-partial reduction record (reference include/lpbf_step_const.hpp:36-37):
 * peak temperature, its row-major index (smallest index wins ties) and the
 * first/last row holding a cell above the melting point.
 * Identity element: {tmax=-1, arg=nx*ny, ymin=ny, ymax=-1}.  24 bytes. */
typedef struct Partial {
    double    tmax;
    long long arg;
    int       ymin;
    int       ymax;
} Partial;

/* One entry of a per-step laser plan: the position the source sits at while
 * step s is computed (the pre-advance position of the reference laser law,
 * src/main.cpp:284-286,351-352) and its on/off multiplier.  Kept in FP64 on
 * the device.  24 bytes. */
typedef struct LaserTick {
    double x_l, y_l, q_scale;
} LaserTick;

/* Fused per-step statistics of the new field T' produced from T (no reference
 * counterpart beyond the Partial prefix):
 *   tmax/arg/ymin/ymax  as in Partial
 *   molten    number of cells with T' > Tm (melt-pool area = molten*dx*dy)
 *   sum_dT    sum over updated cells of (T' - T)          [K]
 *   sum_src   sum over updated cells of dt*S              [K]
 *   sum_loss  sum over updated cells of dt*loss           [K]
 * Multiplying a sum by rho*cp*h*dx*dy gives joules.  Identity element:
 * Partial identity, molten=0 and zero sums.  56 bytes. */
typedef struct StepStats {
    double    tmax;
    long long arg;
    int       ymin;
    int       ymax;
    long long molten;
    double    sum_dT;
    double    sum_src;
    double    sum_loss;
} StepStats;

/* Boundary handling of the full-grid kernels. */
enum {
    LPBF_BC_MIRROR    = 0,  /* ghost = centre on all four edges (reference BC) */
    LPBF_BC_DIRICHLET = 1   /* edge cells held at a fixed temperature */
};

/* Physics selector, same integers as the reference's `phys` argument
 * (include/lpbf_kernels.hpp:21-22). */
enum {
    LPBF_PHYS_BASELINE = 0,
    LPBF_PHYS_EXTENDED = 1
};

#ifdef __cplusplus
}  /* extern "C" */
static_assert(sizeof(StepConst) == 216, "StepConst must keep the reference layout");
static_assert(offsetof(StepConst, radiation) == 176, "StepConst layout");
static_assert(offsetof(StepConst, kmult) == 184, "StepConst layout");
static_assert(offsetof(StepConst, x_l) == 192, "StepConst layout");
static_assert(offsetof(StepConst, q_scale) == 208, "StepConst layout");
static_assert(sizeof(Partial) == 24, "Partial must keep the reference layout");
static_assert(sizeof(LaserTick) == 24, "LaserTick layout");
static_assert(sizeof(StepStats) == 56, "StepStats layout");
#else
_Static_assert(sizeof(StepConst) == 216, "StepConst must keep the reference layout");
_Static_assert(offsetof(StepConst, radiation) == 176, "StepConst layout");
_Static_assert(offsetof(StepConst, kmult) == 184, "StepConst layout");
_Static_assert(offsetof(StepConst, x_l) == 192, "StepConst layout");
_Static_assert(offsetof(StepConst, q_scale) == 208, "StepConst layout");
_Static_assert(sizeof(Partial) == 24, "Partial must keep the reference layout");
_Static_assert(sizeof(LaserTick) == 24, "LaserTick layout");
_Static_assert(sizeof(StepStats) == 56, "StepStats layout");
#endif

#endif /* LPBF_STEP_CONST_H */
