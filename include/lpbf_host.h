/* SPDX-License-Identifier: 0BSD
 *
 * Host-side C ABI of libhedgehog_lpbf_amd.so: configuration arithmetic and
 * laser plans, so callers in any language can build the StepConst and the
 * per-step LaserTick plan the kernels consume.  No device memory is touched.
 */
#ifndef LPBF_HOST_H
#define LPBF_HOST_H

#include "lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mat = {k, rho, cp, Tm, eps, eta}; proc = {P, V, r0, h, hConv, dSub, Tamb,
 * Tsub}.  Same arithmetic as the reference make_const
 * (include/lpbf_step_const.hpp:36-54); kmult = 1 and q_scale = 1. */
void   lpbf_make_const(const double* mat, const double* proc, int radiation,
                       double Lx, double Ly, int nx, int ny, double dt,
                       StepConst* out);

/* Reference dt rules (include/params.hpp:44-47 and :126-131). */
double lpbf_dt_cfl(const double* mat, double Lx, double Ly, int nx, int ny);
double lpbf_dt_cfl_ext(const double* mat, double Lx, double Ly, int nx, int ny,
                       double kmult);

/* Reference single-track law (src/main.cpp:284-286, 351-352) as n_steps
 * LaserTick entries in host memory.  If x_csv / wrapped are non-NULL they
 * receive the post-advance x and wrap flag of each step (metrics.csv and the
 * cruise gate). */
void   lpbf_plan_single_track(double V, double dt, double Lx, double Ly,
                              int n_steps, LaserTick* out, double* x_csv,
                              int* wrapped);

/* Bidirectional raster (new): tracks of `length` along x starting at
 * (x0, y0), alternating direction, `hatch` apart in y, beam off on the
 * jumps and after the last track. */
void   lpbf_plan_raster(double x0, double y0, double length, double hatch,
                        int n_tracks, double V, double dt, int n_steps,
                        LaserTick* out);

#ifdef __cplusplus
}
#endif

#endif /* LPBF_HOST_H */
