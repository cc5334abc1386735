/* SPDX-License-Identifier: 0BSD
 *
 * C ABI of the device-resident run (lpbf::DeviceRun, include/lpbf/
 * device_run.hpp): the reference GpuStepTask + MetricsTask pair
 * (include/gpu_step_task.hpp:135-239, 375-411) behind an opaque handle, for
 * callers that are not C++ (ctypes, cgo, JNI).
 *
 * Every call returns 0 on success or an error code (a hipError_t value or
 * LPBF_ERR_INVALID_ARGUMENT), also recorded for lpbf_get_last_error().
 * lpbf_run_advance returns the number of steps run, or a negative code.
 * Calls on one handle must not overlap (one host thread owns a run).
 */
#ifndef LPBF_RUN_H
#define LPBF_RUN_H

#include "lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpbf_run_options {
    int    precision;     /* 64 (FP64 parity kernels) or 32 (FP32 path) */
    int    bc;            /* LPBF_BC_MIRROR or LPBF_BC_DIRICHLET */
    double T_bc;          /* Dirichlet edge temperature [K] */
    int    batch;         /* steps per host synchronisation, >= 1 */
    int    depth;         /* FP32: steps per launch, 1..4 */
    int    rows;          /* blocked kernels: band height, 0 = built-in */
    double T_ref;         /* FP32: theta origin [K], < 0 = Tamb */
    int    keep_records;  /* 1: keep one lpbf_step_record per step */
    int    graph;         /* FP64: 1 = replay full batches as a captured hipGraph */
    int    f64_depth;     /* FP64 steps per launch, 1..4 (0 = 1); above 1 the
                             temporally blocked kernel, launches ending at cruise
                             and dump steps, no graph; rows = its band height */
    int    cruise_cols;   /* FP32: 1 = launches also end at every cruise step, so
                             each cruise step gets its hot column (W_sub) as in
                             the FP64 path; 0 = only at launch ends */
    int    pad_;
} lpbf_run_options;

/* Per-step result: the merged statistics of the field after the step, the
 * melt-pool width from the molten rows and the sub-grid width through the
 * hottest column (0 when no column was taken). */
typedef struct lpbf_step_record {
    StepStats stats;
    double    W_um;
    double    W_sub_um;
    int       step;
    int       has_col;
} lpbf_step_record;

typedef struct lpbf_run lpbf_run;

/* Laser plan of n_steps ticks in the reference Tick sense
 * (include/gpu_messages.hpp:18-27): plan[s] is the source during step s;
 * x_csv[s] / wrapped[s] the post-advance position and wrap state (cruise
 * gate); dump[s] marks snapshot steps.  x_csv, wrapped, dump may be NULL
 * (no cruise steps, no dumps).  k holds the grid, physics and kmult;
 * Lx, Ly the plate size (cruise gate, W = rows * Ly/(ny-1)); phys is
 * LPBF_PHYS_*; the field starts uniform at Tamb. */
int  lpbf_run_create(const StepConst* k, int phys, double Tm, double Lx, double Ly, double Tamb,
                     const LaserTick* plan, const double* x_csv,
                     const unsigned char* wrapped, const unsigned char* dump,
                     int n_steps, const lpbf_run_options* opt, lpbf_run** out);
void lpbf_run_destroy(lpbf_run* run);

int  lpbf_run_upload(lpbf_run* run, const double* T);            /* nx*ny FP64 */
int  lpbf_run_noise(lpbf_run* run, unsigned long long seed, double base, double amp);
int  lpbf_run_advance(lpbf_run* run, int n_steps);
int  lpbf_run_download(lpbf_run* run, double* T);                /* nx*ny FP64 */

/* out[0..2] = peak T [K], peak W [um], steady W_sub [um]; *steps_done. */
int  lpbf_run_summary(const lpbf_run* run, double* out3, int* steps_done);

/* Number of batches that ran as a replayed hipGraph so far. */
int  lpbf_run_graph_launches(const lpbf_run* run);

/* Copies records [first, first+n) (requires keep_records). */
int  lpbf_run_records(const lpbf_run* run, int first, int n, lpbf_step_record* out);

/* ---- Ensemble: many runs of one plate, one launch per step for all ------
 *
 * The reference's manifest sweep (src/gpu/hh_gpu_main.cpp:234-290: one
 * run_one_row per CSV row, each a GpuStepTask + MetricsTask pair on the same
 * Domain) as one handle: n_runs FP64 runs that share the grid, physics,
 * boundary condition and Tm, and differ in StepConst (k[m]: P, V, r0, eta,
 * dSub, kmult, dt, ...), laser plan, length and initial field.  Every step
 * of every run still running is one lpbf_launch_step_ensemble per step for
 * all runs; the fields, per-step records and summaries equal those of
 * lpbf_run_* runs of the same inputs bit for bit, except the last bits of
 * the FP64 energy sums (bands of another height).
 *
 * plan / x_csv / wrapped hold the runs' ticks back to back (run m's start
 * after the n_steps of runs 0..m-1), with the meaning of lpbf_run_create's;
 * Tamb[m] is run m's initial uniform field (NULL: 300 K).  opt: precision
 * 64 uses bc, T_bc, batch, rows (band height of the ensemble launch,
 * 0..1024, 0 = lpbf_step_ensemble_rows), keep_records and graph.  Precision
 * 32 (baseline physics) steps the runs still running as one lpbf_f32_batch
 * (include/lpbf_batch.h) with up to `depth` steps per launch; launches end
 * at batch ends, at any run's last step and, with cruise_cols, at any run's
 * cruise step; W_sub is taken at launch ends; rows is the FP32 band height
 * (0 = built-in), T_ref the theta origin (< 0: each run's Tamb); graph is
 * ignored.  Its fields equal FP32 lpbf_run_* fields bit for bit.
 * Snapshots are not taken.  Calls on one handle must not overlap. */
typedef struct lpbf_ensemble lpbf_ensemble;
int  lpbf_ensemble_create(const StepConst* k, int n_runs, int phys, double Tm, double Lx,
                          double Ly, const double* Tamb, const LaserTick* plan,
                          const double* x_csv, const unsigned char* wrapped,
                          const int* n_steps, const lpbf_run_options* opt,
                          lpbf_ensemble** out);
void lpbf_ensemble_destroy(lpbf_ensemble* e);
int  lpbf_ensemble_upload(lpbf_ensemble* e, int m, const double* T);   /* before or between advances */
/* Advances the shared step counter by up to n_steps (runs past their own
 * length sit out); returns the steps advanced or a negative code. */
int  lpbf_ensemble_advance(lpbf_ensemble* e, int n_steps);
int  lpbf_ensemble_download(lpbf_ensemble* e, int m, double* T);
/* Run m: out[0..2] = peak T, peak W, steady W_sub; *steps_done of run m. */
int  lpbf_ensemble_summary(const lpbf_ensemble* e, int m, double* out3, int* steps_done);
int  lpbf_ensemble_graph_launches(const lpbf_ensemble* e);
int  lpbf_ensemble_records(const lpbf_ensemble* e, int m, int first, int n,
                           lpbf_step_record* out);

#ifdef __cplusplus
}
static_assert(sizeof(lpbf_run_options) == 64, "lpbf_run_options layout");
static_assert(sizeof(lpbf_step_record) == 80, "lpbf_step_record layout");
#else
_Static_assert(sizeof(lpbf_run_options) == 64, "lpbf_run_options layout");
_Static_assert(sizeof(lpbf_step_record) == 80, "lpbf_step_record layout");
#endif

#endif /* LPBF_RUN_H */
