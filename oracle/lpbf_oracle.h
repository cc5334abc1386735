/* SPDX-License-Identifier: 0BSD
 *
 * CPU oracle for the z-averaged 2D explicit finite-difference heat step of
 * hedgehog_lpbf.  TEST INFRASTRUCTURE ONLY: the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  No
 * product code path links it.
 *
 * Every function restates the reference operation for operation in IEEE
 * double; each cites the reference file:line it follows.  Build with
 * -ffp-contract=off and without -march=native so no FMA is formed, which is
 * how the reference CPU solver was compiled (CMakeLists.txt:64).
 */
#ifndef LPBF_ORACLE_H
#define LPBF_ORACLE_H

#include <stdint.h>
#include "../include/lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields and defaults as the reference include/params.hpp:13-40. */
typedef struct OraMaterial {
    double k, rho, cp, Tm, eps, eta;
} OraMaterial;

typedef struct OraProcess {
    double P, V, r0, h, hConv, dSub, Tamb, Tsub;
    bool   radiation;
} OraProcess;

typedef struct OraDomain {
    double Lx, Ly;
    int    nx, ny;
} OraDomain;

/* One step of the reference tick plan (include/gpu_messages.hpp:13-21). */
typedef struct OraTick {
    int    step;
    double t;
    double x_l, y_l, q_scale;
    double x_laser_csv;  /* post-advance x */
    int    wrapped;      /* post-advance wrap state */
    int    dump;
} OraTick;

/* measure() result (src/main.cpp:158-162 plus x_hot). */
typedef struct OraMetrics {
    double peak_T, W_um, W_sub_um;
    int    x_hot;
} OraMetrics;

/* Parameters of the bidirectional raster path (new, no reference). */
typedef struct OraRaster {
    double x0, y0;       /* start of track 0 [m] */
    double length;       /* track length [m] */
    double hatch;        /* track spacing in y [m] */
    int    n_tracks;
    double V;            /* scan speed [m/s], also used for beam-off jumps */
} OraRaster;

void   ora_defaults(OraMaterial* m, OraProcess* p, OraDomain* d);
int    ora_apply_case(const char* id, OraProcess* p);
double ora_alpha(const OraMaterial* m);
double ora_dx(const OraDomain* d);
double ora_dy(const OraDomain* d);
double ora_dt_cfl(const OraMaterial* m, const OraDomain* d);

/* IN718 extended physics (params.hpp:61-132). */
double ora_interp(double T, const double* xs, const double* ys, int n);
double ora_k_of(double T);
double ora_cp_of(double T);
double ora_k_enh(double T, double kmult);
double ora_k_eff(double T, double kmult);
double ora_cp_eff(double T);
double ora_alpha_eff_max(double rho, double kmult);
double ora_dt_cfl_ext(const OraMaterial* m, const OraDomain* d, double kmult);

void   ora_make_const(const OraMaterial* m, const OraProcess* p,
                      const OraDomain* d, double dt, StepConst* out);

/* Full-grid step with ghost = centre on every edge (lpbf_cuda.cu:154-210). */
void   ora_step(const double* Tin, double* Tout, const StepConst* k, int phys);
/* ora_step on rows [y0, y1) only (mirror edges); used to time the oracle on
 * several host threads. */
void   ora_step_rows(const double* Tin, double* Tout, const StepConst* k, int phys,
                     int y0, int y1);
/* The baseline step of ora_step_rows in FP32 storage and arithmetic (the
 * FP32 CPU baseline; mirror edges). */
void   ora_step_rows_f32(const float* Tin, float* Tout, const StepConst* k, int y0, int y1);
/* ora_step_tile in FP32 (baseline physics): same tile shape in float. */
void   ora_step_tile_f32(const float* tile, float* out, int owned_rows, int y0,
                         const StepConst* k);

/* Row-strip tile step (step_task.hpp:54-206).  `tile` holds rows+2 rows of
 * nx values: halo row 0, owned rows 1..rows, halo row rows+1.  Owned rows
 * start at global row y0.  `out` receives the same shape with the halo rows
 * copied through. */
void   ora_step_tile(const double* tile, double* out, int owned_rows, int y0,
                     const StepConst* k, int phys);

/* Tile bounds and mirrored halo build (main.cpp:115-155). */
void   ora_tile_bounds(int ny, int ntiles, int* y0, int* y1);
void   ora_make_tile(const double* T, int nx, int ny, int y0, int y1,
                     double* tile);

/* One step done the CPU solver's way: build tiles, step each, stitch. */
void   ora_step_tiled(const double* Tin, double* Tout, const StepConst* k,
                      int phys, int ntiles);

/* Global metrics (main.cpp:164-211). */
void   ora_measure(const double* T, int nx, int ny, double dy, double Tm,
                   OraMetrics* out);

/* Sub-grid width from one column (gpu_step_task.hpp:380-395). */
double ora_wsub_from_col(const double* col, int ny, double dy, double Tm);

/* Device reduction semantics (kernels.cu:205-262): grid-stride partials
 * over `nblocks` blocks of 256 threads, then the serial finalize. */
void   ora_reduce_partials(const double* T, int nx, int ny, double Tm,
                           int nblocks, Partial* out);
void   ora_reduce_finalize(const Partial* parts, int nblocks, int nx, int ny,
                           Partial* out);

/* Laser plan (hh_gpu_main.cpp:141-157). */
void   ora_build_plan(double V, double dt, double Lx, double Ly, int n_steps,
                      int snap_every, OraTick* plan);

/* Full single-track run (main.cpp:316-402, lpbf_cuda.cu:423-448).
 * T (nx*ny) is the initial field on entry and the final field on exit.
 * Per-step outputs (each may be NULL): peak_T, W_um, W_sub_um of the field
 * after step s.  Summary out: [peak_T, peak_W, steady_W_sub].
 * ntiles <= 0 steps the full grid; > 0 steps it as row-strip tiles. */
void   ora_run(const OraMaterial* m, const OraProcess* p, const OraDomain* d,
               double dt, int n_steps, int phys, double kmult, int ntiles,
               double* T, double* step_peak, double* step_W,
               double* step_Wsub, double* summary);

/* ---- additions with no reference counterpart --------------------------- */

/* Generalised full-grid step: BC mirror (identical to ora_step) or Dirichlet
 * (edge cells held at T_bc).  If stats != NULL it receives the StepStats of
 * the new field (sums in row-major order). */
void   ora_step_ex(const double* Tin, double* Tout, const StepConst* k,
                   int phys, int bc, double T_bc, double Tm, StepStats* stats);

/* Counter-based uniform noise: value for cell `idx` under `seed`, in
 * [-0.5, 0.5) with 24 random bits (exact in FP32 and FP64). */
double ora_noise(uint64_t seed, uint64_t idx);
void   ora_noise_field(uint64_t seed, double base, double amp, double* T,
                       long long n);

/* Bidirectional raster plan: positions by FP64 running sum of V*dt along
 * the path; q_scale = 0 on the beam-off hatch jumps. */
void   ora_raster_plan(const OraRaster* r, double dt, int n_steps,
                       LaserTick* plan);

/* Reference single-track law as a LaserTick plan. */
void   ora_track_plan(double V, double dt, double Lx, double Ly, int n_steps,
                      LaserTick* plan);

/* Run `n_steps` generalised steps following `plan`; per-step stats into
 * `stats` (n_steps entries, may be NULL).  T in/out. */
void   ora_run_plan(const StepConst* k0, int phys, int bc, double T_bc,
                    double Tm, const LaserTick* plan, int n_steps, double* T,
                    StepStats* stats);

/* The reference CPU solver's time loop with its step-task graph
 * (src/main.cpp:316-348, include/step_task.hpp:54-206): per step the
 * calling thread builds `ntiles` row-strip tiles with mirrored halos (a new
 * buffer each, main.cpp:125-155), `nthreads` worker threads execute them
 * (each writes a new buffer and drops the input, step_task.hpp:78-84,
 * 190-191), and the calling thread stitches the owned rows back as the
 * tiles complete, in tile order.  prec 64: T is double, phys 0/1; prec 32:
 * T is float, baseline physics (ora_step_tile_f32).  The laser of step s is
 * plan[s].  T (nx*ny, row-major) is the initial field on entry and the
 * final field on exit.  phase_s (may be NULL) receives the seconds spent
 * building tiles, waiting for and stitching results, and in total.
 * Returns 0, or -1 for invalid arguments or a failed allocation. */
int    ora_steptask_run(int prec, const StepConst* k, int phys, const LaserTick* plan,
                        int n_steps, int ntiles, int nthreads, void* T, double* phase_s);

#ifdef __cplusplus
}
#endif

#endif /* LPBF_ORACLE_H */
