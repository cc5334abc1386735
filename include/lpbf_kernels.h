/* SPDX-License-Identifier: 0BSD
 *
 * C ABI of the MI355X (gfx950) step library libhedgehog_lpbf_amd.so.
 *
 * The first block keeps the names, argument order, argument meaning and
 * calling convention of the reference launch seam (hedgehog_lpbf
 * include/lpbf_kernels.hpp:19-55, defined at src/gpu/kernels.cu:280-319), with
 * hipStream_t in place of cudaStream_t:
 *   - the caller allocates and frees every device buffer;
 *   - Tin and Tout never alias;
 *   - *k is copied into the launch, so the caller may change it on return;
 *   - every call is asynchronous on `stream`, returns nothing and never
 *     synchronises; a launch failure or invalid argument is reported through
 *     lpbf_get_last_error() (and the HIP runtime's hipGetLastError()).
 *
 * Fields are FP64 row-major nx*ny arrays (index y*nx + x) unless stated.
 */
#ifndef LPBF_KERNELS_H
#define LPBF_KERNELS_H

#include <hip/hip_runtime_api.h>
#include "lpbf_step_const.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference seam ------------------------------------------------------ */

/* One explicit-Euler step of the full grid, ghost = centre on all edges.
 * phys 0 baseline, 1 extended (IN718 tables).
 * Replaces lpbf_launch_step, include/lpbf_kernels.hpp:24-25. */
void lpbf_launch_step(const double* Tin, double* Tout, const StepConst* k,
                      int phys, hipStream_t stream);

/* Grid-stride block partials of T: `nblocks` records into `out`, each the
 * same record the reference reduce_kernel writes for that block.
 * Replaces lpbf_launch_reduce, include/lpbf_kernels.hpp:30-31. */
void lpbf_launch_reduce(const double* T, int nx, int ny, double Tm,
                        Partial* out, int nblocks, hipStream_t stream);

/* Step fused with the partial reduction of the new field.  Writes exactly
 * ((nx+15)/16)*((ny+15)/16) records into `out` (records this kernel does not
 * need hold the identity), so finalizing them with that count gives the
 * global Partial of Tout.
 * Replaces lpbf_launch_step_reduce, include/lpbf_kernels.hpp:37-39. */
void lpbf_launch_step_reduce(const double* Tin, double* Tout,
                             const StepConst* k, int phys, double Tm,
                             Partial* out, hipStream_t stream);

/* Merge `nblocks` partials into out[0] on the device (max value, smallest
 * index on ties; min ymin; max ymax).
 * Replaces lpbf_launch_reduce_finalize, include/lpbf_kernels.hpp:44-46. */
void lpbf_launch_reduce_finalize(const Partial* parts, int nblocks, int nx,
                                 int ny, Partial* out, hipStream_t stream);

/* Copy column x_hot = glob->arg % nx of T into out_col[0..ny).
 * Replaces lpbf_launch_extract_col, include/lpbf_kernels.hpp:51-53. */
void lpbf_launch_extract_col(const double* T, int nx, int ny,
                             const Partial* glob, double* out_col,
                             hipStream_t stream);

/* ---- status (new) ---------------------------------------------------------- */

/* First error recorded by any lpbf_* call since the last query (a hipError_t
 * value, or LPBF_ERR_INVALID_ARGUMENT), then cleared.  0 = success. */
#define LPBF_ERR_INVALID_ARGUMENT 1000
#define LPBF_ERR_SCRATCH_LIMIT    1001   /* lpbf_launch_step_f32_noalloc: workspace too large */
int         lpbf_get_last_error(void);
const char* lpbf_error_string(int err);

/* ---- FP64 additions (no reference counterpart) ------------------------------ */

/* The reference StepTileTask::execute (include/step_task.hpp:54-206) on a
 * device-resident row-strip tile: tile_in holds owned_rows+2 rows of k->nx
 * values (halo row, owned rows, halo row; halos carry the mirror at the
 * domain edges); the owned rows of tile_out receive the update, its halo rows
 * a copy of tile_in's.  y0 is the global row of the first owned row. */
void lpbf_launch_step_tile(const double* tile_in, double* tile_out,
                           int owned_rows, int y0, const StepConst* k,
                           int phys, hipStream_t stream);

/* Generalised full-grid step: bc LPBF_BC_MIRROR (bitwise lpbf_launch_step) or
 * LPBF_BC_DIRICHLET (edge cells set to T_bc, interior uses all four
 * neighbours).  If block_stats != NULL it receives
 * lpbf_step_ex_nblocks(nx, ny) block records of the new field, which
 * lpbf_launch_stats_finalize merges into one StepStats. */
void lpbf_launch_step_ex(const double* Tin, double* Tout, const StepConst* k,
                         int phys, int bc, double T_bc, double Tm,
                         StepStats* block_stats, hipStream_t stream);
int  lpbf_step_ex_nblocks(int nx, int ny);

/* Plan-driven step for launch sequences that are captured once and
 * replayed (hipGraph): the step index lives in device memory.  One launch
 * does lpbf_launch_step_ex for step s = *d_step with the laser taken from
 * d_plan[s] (k's x_l, y_l, q_scale ignored), writes its block records to
 * block_stats (lpbf_step_ex_plan_nblocks(nx, ny) of them), and the last
 * workgroup to finish merges them into ring[s % ring_len], copies column
 * ring[..].arg % nx of Tout into cols[(s % ring_len) * ny ...] when cols and
 * d_cruise are given and d_cruise[s] != 0, resets *d_ticket to 0 and sets
 * *d_step = s + 1.  *d_ticket must be 0 before the first launch. */
int  lpbf_step_ex_plan_nblocks(int nx, int ny);
void lpbf_launch_step_ex_plan(const double* Tin, double* Tout, const StepConst* k,
                              int phys, int bc, double T_bc, double Tm,
                              const LaserTick* d_plan, int* d_step,
                              const unsigned char* d_cruise, StepStats* block_stats,
                              int* d_ticket, StepStats* ring, int ring_len, double* cols,
                              hipStream_t stream);

/* Batch form of the plan-driven step, for a captured sequence of launches
 * j = 0..n-1 (every node carries its own j, so there is no device counter
 * and no atomic): the caller copies the batch's laser ticks and cruise flags
 * to d_ticks[0..n) and d_cruise[0..n) (device) before each replay.  Launch j
 * steps Tin -> Tout with the laser of d_ticks[j] and writes its records
 * (one per wavefront) into block_stats[(j % 2) * nblk ...] (nblk =
 * lpbf_step_ex_batch_nblocks(nx, ny); block_stats holds 2 * nblk records);
 * one extra workgroup meanwhile finishes step j-1: it merges that step's
 * records into ring[j-1] and, when d_cruise[j-1] != 0, copies the column of
 * its maximum from Tin (step j-1's output) into cols[(j-1)*ny ...].
 * lpbf_launch_step_ex_batch_flush(T_n, ..., j = n, ...) then finishes step
 * n-1 from its output field.  Fields and columns equal those of
 * lpbf_launch_step_ex + lpbf_launch_stats_finalize + lpbf_launch_extract_col
 * bit for bit, the ring's max, index, rows and molten count too; its FP64
 * sums agree to rounding (the block decomposition differs). */
int  lpbf_step_ex_batch_nblocks(int nx, int ny);
void lpbf_launch_step_ex_batch(const double* Tin, double* Tout, const StepConst* k,
                               int phys, int bc, double T_bc, double Tm,
                               const LaserTick* d_ticks, int j, const unsigned char* d_cruise,
                               StepStats* block_stats, StepStats* ring, double* cols,
                               hipStream_t stream);
void lpbf_launch_step_ex_batch_flush(const double* T, int nx, int ny, int j,
                                     const unsigned char* d_cruise,
                                     const StepStats* block_stats, StepStats* ring,
                                     double* cols, hipStream_t stream);

/* Ensemble form of the batch step: many runs of one plate (the reference's
 * manifest sweeps, src/gpu/hh_gpu_main.cpp:234-290: same Domain, rows that
 * differ in process parameters) advanced by one launch per batch step.  Each
 * run has its own constants, two field buffers (step s reads buf[s & 1] and
 * writes the other, so after n steps the field is in buf[n & 1]), whole-run
 * laser plan and cruise flags, 2 * nblk block records (nblk =
 * lpbf_step_ensemble_nblocks(nx, ny, rows)), and a ring and column buffer
 * for one batch.  Launch j computes step s = *d_s0 + j of every run still
 * running (*d_s0 = the batch's first step, in device memory, set by the
 * caller before the batch) and finishes step s - 1 into ring[j-1] and, on
 * cruise steps, cols[(j-1)*col_pitch ...] (ny values), exactly as
 * lpbf_launch_step_ex_batch; runs whose n_steps are reached sit out.  The
 * flush (j = batch length) finishes the batch's last step.  rows is the
 * band height of a wavefront (1..1024; 0 = lpbf_launch_step_ex_batch's, so
 * the block records and their sums equal a single run's); phys, bc, T_bc
 * and Tm are shared. */
typedef struct lpbf_ensemble_run {
    StepConst            k;        /* the run's constants; x_l, y_l, q_scale unused */
    double*              buf[2];   /* field buffers (device) */
    const LaserTick*     plan;     /* laser per step (device), n_steps ticks */
    const unsigned char* cruise;   /* cruise flag per step (device), n_steps bytes */
    StepStats*           rec;      /* 2 * nblk block records (device) */
    StepStats*           ring;     /* per batch step (device) */
    double*              cols;     /* per batch step, ny values each (device) */
    int                  n_steps;
    int                  col_pitch; /* doubles from one step's column to the next, >= ny */
} lpbf_ensemble_run;
int  lpbf_step_ensemble_nblocks(int nx, int ny, int rows);   /* 0: invalid */
/* Band height for n_runs runs of an nx x ny plate (what lpbf::Ensemble uses
 * by default); 0 for an invalid shape. */
int  lpbf_step_ensemble_rows(int nx, int ny, int n_runs);
void lpbf_launch_step_ensemble(const lpbf_ensemble_run* d_runs, int n_runs, int nx, int ny,
                               int rows, int phys, int bc, double T_bc, double Tm,
                               const int* d_s0, int j, hipStream_t stream);
void lpbf_launch_step_ensemble_flush(const lpbf_ensemble_run* d_runs, int n_runs, int nx,
                                     int ny, int rows, const int* d_s0, int j,
                                     hipStream_t stream);

/* Merge block StepStats in a fixed order (deterministic FP64 sums). */
void lpbf_launch_stats_finalize(const StepStats* parts, int nblocks, int nx,
                                int ny, StepStats* out, hipStream_t stream);

/* FP64 steps with temporal blocking: one launch advances Tin by `depth`
 * (1..4) steps into Tout, the laser of step step0+j taken from d_plan[step0+j]
 * (device, d_plan must hold step0+depth entries; k's x_l, y_l, q_scale are
 * ignored), bc LPBF_BC_MIRROR or LPBF_BC_DIRICHLET as lpbf_launch_step_ex.
 * Every cell of every step evaluates the same expression as
 * lpbf_launch_step_ex, so Tout is bitwise the field of `depth` such calls.
 * rows: band height (0 = built-in).  If stats != NULL it receives `depth`
 * StepStats (one per step, merged on the device; the FP64 sums are added in
 * a different order than lpbf_launch_step_ex's, so they agree to rounding)
 * and work must hold lpbf_f64_blocked_workspace_bytes(nx, ny, D, rows)
 * bytes of device memory, which covers every launch of depth <= D with the
 * same rows. */
size_t lpbf_f64_blocked_workspace_bytes(int nx, int ny, int depth, int rows);
void lpbf_launch_step_f64_blocked(const double* Tin, double* Tout, const StepConst* k,
                                  int phys, int bc, double T_bc, double Tm,
                                  const LaserTick* d_plan, int step0, int depth, int rows,
                                  StepStats* stats, void* work, hipStream_t stream);

/* Autotuner for lpbf_launch_step_f64_blocked: times depths 1..max_depth and
 * a set of band heights on scratch copies of a k->nx x k->ny field
 * (allocated and freed here, synchronising `stream`) and returns the fastest
 * per step.  d_plan needs max_depth entries.  Returns 0 or an error code
 * (also recorded for lpbf_get_last_error). */
int lpbf_f64_blocked_autotune(const StepConst* k, int phys, int bc, double T_bc, double Tm,
                              const LaserTick* d_plan, int n_plan, int max_depth,
                              hipStream_t stream, int* best_depth, int* best_rows,
                              double* best_ms_per_step);

/* Counter-based noise field: T[i] = base + 2*amp*u(seed, i), u in [-0.5,0.5)
 * from SplitMix64 with 24 random bits; n values. */
void lpbf_launch_noise_f64(double* T, long long n, unsigned long long seed,
                           double base, double amp, hipStream_t stream);

/* y[i] = exp(x[i]) with the device exp the FP64 step kernels call, n values.
 * No reference counterpart: the reference step calls libm exp on the CPU
 * (include/step_task.hpp:113, 168) and the CUDA device exp on the GPU
 * (src/gpu/kernels.cu:115); this exposes the one libm function at the
 * parity boundary so its distance to a host libm can be measured. */
void lpbf_probe_exp_f64(const double* x, double* y, long long n, hipStream_t stream);

/* ---- FP32 path (no reference counterpart) ---------------------------------
 *
 * Field layout: theta = T - T_ref in FP32, row y at theta + y*ld, ld >= nx,
 * ld % 4 == 0, base 16-byte aligned.  Columns [nx, ld) of every row are
 * scratch: the step may overwrite them in theta_out.
 *
 * One launch advances `depth` (1..4) steps, theta_in -> theta_out, reading
 * the laser for step step0+j from d_plan[step0+j] (device, FP64).  The grid
 * is k->nx x k->ny with spacing and physics from *k (baseline physics; k's
 * x_l, y_l, q_scale are ignored).  bc: LPBF_BC_MIRROR or LPBF_BC_DIRICHLET
 * (edges held at T_bc).  `rows` is the band height of a wavefront tile
 * (0 = built-in choice).  If d_stats != NULL it receives `depth` StepStats,
 * one per step, of the fields after each step.  d_work must hold
 * lpbf_f32_workspace_bytes(nx, ny, D, rows) bytes of device memory, which
 * covers every launch of depth <= D with the same rows.
 */
size_t lpbf_f32_workspace_bytes(int nx, int ny, int depth, int rows);
void   lpbf_launch_step_f32(const float* theta_in, float* theta_out, int ld,
                            const StepConst* k, double T_ref, int bc,
                            double T_bc, double Tm, const LaserTick* d_plan,
                            int step0, int depth, int rows, StepStats* d_stats,
                            void* d_work, hipStream_t stream);

/* FP32 form of the extended physics (the reference's phys = 1 update,
 * include/step_task.hpp:132-187 and src/gpu/kernels.cu:125-147:
 * temperature-dependent IN718 conductivity and heat capacity, latent heat
 * over the mushy interval, k->kmult melt-pool conductivity enhancement) on
 * the pitched FP32 theta field of lpbf_launch_step_f32: one step
 * theta_in -> theta_out with the laser of d_plan[step], and if d_stats !=
 * NULL the step's merged StepStats (sum_dT: the summed temperature change;
 * sum_src, sum_loss: dt*q_in/(rho cp_eff h) and dt*(q_top+q_bot)/(rho
 * cp_eff h), in K).  bc as for lpbf_launch_step_f32.  d_work must hold
 * lpbf_f32_ext_workspace_bytes(nx, ny) bytes.  Its error against the FP64
 * path is stated where it is tested (tests/test_gpu_f32_ext.py). */
size_t lpbf_f32_ext_workspace_bytes(int nx, int ny);
void   lpbf_launch_step_f32_ext(const float* theta_in, float* theta_out, int ld,
                                const StepConst* k, double T_ref, int bc, double T_bc,
                                double Tm, const LaserTick* d_plan, int step,
                                StepStats* d_stats, void* d_work, hipStream_t stream);

/* lpbf_launch_step_f32 without a caller workspace, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * per-workgroup statistics records live in a fixed device array of the
 * library, LPBF_F32_NOALLOC_SCRATCH_BYTES per device, held until the process
 * ends; no call allocates or frees memory.  Same arguments, kernels and
 * results as lpbf_launch_step_f32.  A launch whose records
 * (depth * workgroups * sizeof(StepStats)) exceed the array is refused with
 * LPBF_ERR_SCRATCH_LIMIT (the 8192^2 plate at depth 4 needs ~0.35 MB with
 * the built-in band height).  Every call on a device shares that device's
 * array: calls that overlap on one device (other streams, other host
 * threads) race; calls in order on one stream do not. */
#define LPBF_F32_NOALLOC_SCRATCH_BYTES (4u << 20)
void   lpbf_launch_step_f32_noalloc(const float* theta_in, float* theta_out, int ld,
                                    const StepConst* k, double T_ref, int bc, double T_bc,
                                    double Tm, const LaserTick* d_plan, int step0, int depth,
                                    int rows, StepStats* d_stats, hipStream_t stream);

/* The per-step merge that lpbf_launch_step_f32 runs when d_stats != NULL,
 * as its own call: after a step launch with d_stats == NULL and the same
 * (nx, ny, depth, rows, d_work), writes the `depth` StepStats to d_stats. */
void   lpbf_launch_f32_stats_finalize(const void* d_work, int nx, int ny,
                                      int depth, int rows, StepStats* d_stats,
                                      hipStream_t stream);

/* Autotuner: times lpbf_launch_step_f32 on scratch copies of a k->nx x
 * k->ny field (allocated and freed here, synchronising `stream`) for
 * temporal depths 2..max_depth (1 if max_depth == 1) and a set of band
 * heights, and returns the fastest per step.  d_plan needs max_depth entries.
 * Returns 0 or an error code (also recorded for lpbf_get_last_error). */
int    lpbf_f32_autotune(const StepConst* k, double T_ref, int bc, double T_bc,
                         double Tm, const LaserTick* d_plan, int n_plan,
                         int max_depth, hipStream_t stream, int* best_depth,
                         int* best_rows, double* best_ms_per_step);

/* theta = base + 2*amp*u(seed, y*nx+x) (same u as lpbf_launch_noise_f64). */
void   lpbf_launch_noise_f32(float* theta, int nx, int ny, int ld,
                             unsigned long long seed, double base, double amp,
                             hipStream_t stream);

/* FP32 counterpart of lpbf_launch_extract_col: out_col[y] = T_ref +
 * theta[y*ld + x_hot] with x_hot = glob->arg % nx (glob: the step's merged
 * StepStats, whose first fields are a Partial). */
void   lpbf_launch_extract_col_f32(const float* theta, int ld, int nx, int ny,
                                   double T_ref, const StepStats* glob,
                                   double* out_col, hipStream_t stream);

/* Conversions between the FP64 row-major T and the pitched FP32 theta. */
void   lpbf_launch_theta_to_T(const float* theta, int ld, int nx, int ny,
                              double T_ref, double* T, hipStream_t stream);
void   lpbf_launch_T_to_theta(const double* T, int nx, int ny, double T_ref,
                              float* theta, int ld, hipStream_t stream);

#ifdef __cplusplus
}
static_assert(sizeof(lpbf_ensemble_run) == 280, "lpbf_ensemble_run layout");
#else
_Static_assert(sizeof(lpbf_ensemble_run) == 280, "lpbf_ensemble_run layout");
#endif

#endif /* LPBF_KERNELS_H */
