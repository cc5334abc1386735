/* SPDX-License-Identifier: 0BSD
 *
 * Batched FP32 instances (no reference counterpart): many process-parameter
 * instances of uneven grid size advanced together, each launch stepping every
 * instance of a group by `depth` steps with one multi-instance kernel, the
 * whole run capturable as one hipGraph.
 *
 * The reference runs such sweeps one after another (src/gpu/hh_gpu_main.cpp
 * run_manifest, :234-290, one graph and one device field per row); here
 * every instance is resident at once and the launch grid covers all of them.
 *
 * Instances carry a size class; instances of one class share the temporal
 * depth (1..4) and band height, chosen by lpbf_f32_batch_set_tiling or
 * measured by lpbf_f32_batch_autotune.  Classes with equal depth (and equal
 * radiation switch) share a launch.
 *
 * Fields follow the lpbf_launch_step_f32 layout (theta = T - T_ref, pitch ld,
 * ld % 4 == 0, 16-byte aligned).  Step s of instance i uses d_plan[s] of that
 * instance and, if d_stats != NULL, writes d_stats[s].  Each instance
 * alternates between theta0 and theta1; lpbf_f32_batch_current tells which
 * holds its latest field.
 *
 * Return values: 0 or an error code (hipError_t value or
 * LPBF_ERR_INVALID_ARGUMENT), also recorded for lpbf_get_last_error().
 * Calls on one batch must not overlap.
 */
#ifndef LPBF_BATCH_H
#define LPBF_BATCH_H

#include <hip/hip_runtime_api.h>
#include "lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpbf_f32_instance {
    StepConst        k;           /* grid and physics; x_l, y_l, q_scale unused */
    float*           theta0;      /* initial field (device) */
    float*           theta1;      /* second buffer (device) */
    int              ld;          /* row pitch in floats */
    int              size_class;  /* >= 0 */
    double           T_ref;       /* theta origin [K] */
    double           T_bc;        /* Dirichlet edge temperature [K] */
    double           Tm;          /* melting point for the statistics [K] */
    const LaserTick* d_plan;      /* laser per step (device), n_steps ticks */
    StepStats*       d_stats;     /* per-step statistics (device), n_steps records, or NULL */
    int              n_steps;     /* length of d_plan (and d_stats) */
    int              pad_;
} lpbf_f32_instance;

typedef struct lpbf_f32_batch lpbf_f32_batch;

/* bc: LPBF_BC_MIRROR or LPBF_BC_DIRICHLET for every instance.  The batch
 * keeps a copy of the descriptors; the device buffers stay the caller's. */
int  lpbf_f32_batch_create(const lpbf_f32_instance* inst, int n_inst, int bc,
                           lpbf_f32_batch** out);
void lpbf_f32_batch_destroy(lpbf_f32_batch* b);

/* Tiling of one size class: depth 1..4, rows = band height (0 = built-in).
 * Changing the tiling (or autotuning) discards a captured graph. */
int  lpbf_f32_batch_set_tiling(lpbf_f32_batch* b, int size_class, int depth, int rows);
int  lpbf_f32_batch_get_tiling(const lpbf_f32_batch* b, int size_class, int* depth,
                               int* rows, int* n_members);

/* Picks each class's depth and band height by timing, with the instances'
 * own plans from step 0: (1) every class alone, for depths 2..max_depth (1
 * if max_depth == 1) and a set of band heights; (2) schedules: every class
 * at its own best depth, or every class at one depth (so that they share
 * one launch), each timed as its launches; (3) in each shared launch, each
 * class's band height again, timing the shared launch.  Timings within 2%
 * count as equal (then the taller bands, or the schedule with fewer
 * launches, are kept).  The fields are saved and restored (synchronises
 * `stream`); *ms_per_step (may be NULL) receives the picked schedule's time
 * per step, its launches timed one by one. */
int  lpbf_f32_batch_autotune(lpbf_f32_batch* b, int max_depth, hipStream_t stream,
                             double* ms_per_step);

/* Enqueues steps [step0, step0+n_steps) of every instance on `stream`
 * (asynchronous, no host synchronisation; may be captured).  The steps must
 * lie within every instance's plan. */
int  lpbf_f32_batch_enqueue(lpbf_f32_batch* b, int step0, int n_steps, hipStream_t stream);

/* Captures the launches of lpbf_f32_batch_enqueue(step0, n_steps) as one
 * hipGraph (nothing runs).  Every group must have an even number of launches
 * so that replays leave each field in the buffer it started in; a schedule
 * with an odd count is made even by splitting its last multi-step launch,
 * and an odd count of single-step launches is refused. */
int  lpbf_f32_batch_capture(lpbf_f32_batch* b, int step0, int n_steps, hipStream_t stream);
/* Launches the captured graph on `stream` (may be repeated).  Refused with
 * LPBF_ERR_INVALID_ARGUMENT when an eager lpbf_f32_batch_enqueue since the
 * capture has left some field in the other buffer than at capture (an odd
 * number of launches); an even number of eager launches restores it. */
int  lpbf_f32_batch_launch_graph(lpbf_f32_batch* b, hipStream_t stream);

/* Enqueues steps [step0, step0+n_steps) of every instance as exactly one
 * launch per launch group (1 <= n_steps <= the depth of every class), with
 * no splitting for graph replay: every field moves to its other buffer.
 * For callers that end launches at steps of their own choosing (the FP32
 * ensemble ends them at batch ends, run ends and cruise steps). */
int  lpbf_f32_batch_enqueue_launch(lpbf_f32_batch* b, int step0, int n_steps,
                                   hipStream_t stream);

/* 0 if theta0, 1 if theta1 holds instance i's latest field; < 0 on error. */
int  lpbf_f32_batch_current(const lpbf_f32_batch* b, int instance);

/* Hot columns of many FP32 fields after one step (the FP32 counterpart of
 * the ensemble's per-step column copy, lpbf_ensemble_run.cols): for each job
 * whose cruise[step] is nonzero and step < n_steps, writes the ny values
 * T_ref + theta[q][y*ld + x_hot], x_hot = stats[step].arg % nx, to
 * col + slot*col_pitch.  One workgroup per job; asynchronous on `stream`. */
typedef struct lpbf_f32_col_job {
    const float*         theta[2];  /* theta[q] holds the field after `step` */
    const StepStats*     stats;     /* merged per-step statistics, by step */
    const unsigned char* cruise;    /* per-step copy flags */
    double*              col;       /* destination of slot 0 */
    double               T_ref;     /* theta origin [K] */
    int                  ld, nx, ny, n_steps;
} lpbf_f32_col_job;

int  lpbf_launch_hot_cols_f32(const lpbf_f32_col_job* d_jobs, int n_jobs, int q, int step,
                              int slot, long long col_pitch, hipStream_t stream);

/* Workgroups and launches per enqueue of n_steps (for traffic accounting). */
int  lpbf_f32_batch_launches(const lpbf_f32_batch* b, int n_steps, int* n_launches,
                             long long* cells_per_step);

#ifdef __cplusplus
}
static_assert(sizeof(lpbf_f32_instance) == 288, "lpbf_f32_instance layout");
static_assert(sizeof(lpbf_f32_col_job) == 64, "lpbf_f32_col_job layout");
#else
_Static_assert(sizeof(lpbf_f32_instance) == 288, "lpbf_f32_instance layout");
_Static_assert(sizeof(lpbf_f32_col_job) == 64, "lpbf_f32_col_job layout");
#endif

#endif /* LPBF_BATCH_H */
