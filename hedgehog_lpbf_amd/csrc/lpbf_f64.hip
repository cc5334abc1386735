// SPDX-License-Identifier: 0BSD
// FP64 step kernels for gfx950: the parity path of the library.
//
// Each cell update evaluates the reference expression in the reference's
// operation order (include/step_task.hpp:93-129 / src/gpu/kernels.cu:106-147)
// with FP contraction off, so every product and sum rounds as written and the
// only difference from the CPU restatement is the device exp().
//
// Memory layout and schedule (MI355X-first):
//   * a wavefront owns a 128-column x R-row block of the field; lane l holds
//     columns c0+l and c0+64+l, so every row load is a contiguous 512-byte
//     wavefront access;
//   * the block is streamed top to bottom keeping the previous, current and
//     next row in registers, so each input row is read from HBM once per
//     block (plus one halo row above and below);
//   * x neighbours come from cross-lane shuffles; lanes 0 and 63 load the
//     one column to the left/right of the block;
//   * four independent wavefronts form a 256-thread workgroup; reductions
//     are wavefront shuffles followed by a 4-way merge in LDS.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "lpbf_common.hpp"
#include "lpbf_f64_cell.hpp"
#include "lpbf_kernels.h"

namespace lpbf_amd {
namespace {

constexpr int kBlockCols = 128;              // columns per wavefront
constexpr int kThreads = 256;                // 4 wavefronts
constexpr int kRowsPerWave = 32;
constexpr int kAhead = 3;                    // rows in flight ahead of the one computed
constexpr int kMaxEnsembleRows = 1024;       // band height limit of the ensemble launch

enum { BC_MIRROR = LPBF_BC_MIRROR, BC_DIRICHLET = LPBF_BC_DIRICHLET, BC_TILE = 2 };
enum { RED_NONE = 0, RED_PARTIAL = 1, RED_STATS = 2, RED_RING = 3, RED_BATCH = 4,
       RED_ENSEMBLE = 5 };

// The rows being updated.  For BC_TILE, in/out point at the first owned row
// and rows -1 and ny (the halos) exist in both buffers.
using gdouble = __attribute__((address_space(1))) double;

struct Region {
    const double* in;
    double* out;
    int nx, ny;     // columns; rows to update
    int y_off;      // global row of region row 0
    int bc;
    double T_bc;
    double Tm;
    int R;          // rows per wavefront block
    int nwx;        // wavefront blocks across
    int ntasks;     // wavefront blocks in total
    int ny_glob;    // global row count (reduction identity)
    // plan-driven launches (graph replay): the laser of step *step comes from
    // plan[*step] instead of k's x_l, y_l, q_scale; with RED_RING the last
    // workgroup to finish (ticket) merges the block records into
    // ring[s % ring_len], copies the hot column on cruise steps, resets the
    // ticket and advances *step.  With RED_BATCH the step is the launch
    // argument j (plan[j], cruise[j]) and an extra last workgroup finishes
    // step j-1 (see k_step_f64).
    const LaserTick* plan;
    int* step;
    int* ticket;
    StepStats* ring;
    int ring_len;
    const unsigned char* cruise;
    double* cols;
    Recips rc;      // the launch's quotients, computed on the host (recips(k))
};

// One run of an ensemble (include/lpbf_kernels.h, lpbf_ensemble_run).
using EnsMember = lpbf_ensemble_run;

// Finishes step j of a batch: merges its nblk block records (in a fixed
// order, so the sums are reproducible) into ring[j] and, on cruise steps,
// copies the column of the step's maximum from `field` (the step's output)
// into cols[j*col_pitch ...] (ny values).  One 256-thread workgroup.
__device__ void finish_batch_step(const StepStats* __restrict__ rec, int nblk, int nx, int ny,
                                  int j, const double* __restrict__ field,
                                  const unsigned char* __restrict__ cruise,
                                  StepStats* __restrict__ ring, double* __restrict__ cols,
                                  long long col_pitch) {
    __shared__ int s_xhot;
    const StepStats st = block_merge_stats(rec, nblk, nx, ny, ring + j);
    if (threadIdx.x == 0) s_xhot = static_cast<int>(st.arg % nx);
    __syncthreads();
    if (cols && cruise && cruise[j]) {
        double* col = cols + j * col_pitch;
        for (int y = threadIdx.x; y < ny; y += kThreads)
            col[y] = field[static_cast<long long>(y) * nx + s_xhot];
    }
}

template <int PHYS, int RED>
__global__ __launch_bounds__(kThreads) void k_step_f64(Region g, StepConst k,
                                                       void* red_out, int red_fill) {
    // bid: this workgroup within its step's grid (ncb compute workgroups,
    // plus the finishing one in the batch forms)
    int bid = blockIdx.x;
    int ncb = static_cast<int>(gridDim.x) - (RED == RED_BATCH ? 1 : 0);
    StepStats* recs = static_cast<StepStats*>(red_out);
    int tick = RED == RED_BATCH ? red_fill : 0;
    if (RED == RED_ENSEMBLE) {
        // run m of the ensemble; its step s = s0 + j (s0 = *g.step, the
        // batch start shared by every run) reads buf[s & 1]
        ncb = (g.ntasks + kThreads / kWave - 1) / (kThreads / kWave);
        const int m = static_cast<int>(blockIdx.x) / (ncb + 1);
        bid = static_cast<int>(blockIdx.x) - m * (ncb + 1);
        const EnsMember& M = static_cast<const EnsMember*>(red_out)[m];
        const int j = red_fill, s = *g.step + j;
        if (bid == ncb) {
            if (j > 0 && s - 1 < M.n_steps)
                finish_batch_step(M.rec + ((s - 1) & 1) * ncb, ncb, g.nx, g.ny, j - 1,
                                  M.buf[s & 1], M.cruise + (s - j), M.ring, M.cols,
                                  M.col_pitch);
            return;
        }
        if (s >= M.n_steps) return;    // this run has ended
        g.in = M.buf[s & 1];
        g.out = M.buf[(s & 1) ^ 1];
        g.plan = M.plan;
        k = M.k;
        recs = M.rec + (s & 1) * ncb;
        tick = s;
    }
    if (RED == RED_BATCH && blockIdx.x == gridDim.x - 1) {
        // the extra workgroup finishes the previous step of the batch while
        // the others compute this one: its records (one per wavefront) are
        // in the other half of the record buffer and its output is this
        // launch's input
        const int j = red_fill, nblk = ncb * (kThreads / kWave);
        if (j > 0)
            finish_batch_step(recs + ((j - 1) & 1) * nblk, nblk, g.nx, g.ny, j - 1, g.in,
                              g.cruise, g.ring, g.cols, g.ny);
        return;
    }
    if (RED == RED_BATCH) recs += (red_fill & 1) * ncb * (kThreads / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform: the block, its rows and the row loop live in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int task = bid * (kThreads / kWave) + wave;
    const int nx = g.nx, ny = g.ny;

    PartialAcc pa = PartialAcc::identity(nx, g.ny_glob);
    long long molten = 0;
    double s_dT = 0.0, s_src = 0.0, s_loss = 0.0;
    // A lane meets its cells in increasing index order (row by row, x0 then
    // x1), so a strict > keeps the smallest index among equal maxima; the
    // position is kept as 2*row + (second column) and turned into the index
    // once, before the reductions.
    int bpos = -1, xs0 = 0, xs1 = 0;

    if (task < g.ntasks) {
        // the ensemble form takes each run's constants from memory
        const Recips rcp_k = RED == RED_ENSEMBLE ? recips(k) : g.rc;
        const int tx = task % g.nwx, ty = task / g.nwx;
        const int c0 = tx * kBlockCols;
        const int ya = ty * g.R;
        const int yb = min(ya + g.R, ny);
        const int x0 = c0 + lane, x1 = c0 + kWave + lane;
        const bool a0 = x0 < nx, a1 = x1 < nx;
        xs0 = x0;
        xs1 = x1;
        const int ecol = (lane == 0) ? c0 - 1 : c0 + kBlockCols;
        const bool tile = (g.bc == BC_TILE);

        // rows through global-address-space pointers: the ensemble form reads
        // its buffer pointers from memory, where the compiler cannot infer
        // the address space and would use flat accesses, which also count on
        // the scalar-load counter (every wait for a scalar value would drain
        // the row prefetches)
        const gdouble* gin = (const gdouble*)g.in;
        gdouble* gout = (gdouble*)g.out;
        // every row load is unconditional, at a column clamped into the
        // row: a load under a divergent lane condition is waited for before
        // the next one is issued, which serialised the first rows' round
        // trips (the bulk of a small plate's step).  Lanes past the last
        // column compute nothing, and the edge ghost (ce) is used only where
        // it is a real column, so the clamped values are never used.  The
        // row above the first row of a mirror-edged field is that row again
        // (the ghost equals the centre), read rather than copied so that no
        // load waits for another.
        const int lx0 = min(x0, nx - 1), lx1 = min(x1, nx - 1);
        const int lxe = min(max(ecol, 0), nx - 1);
        const gdouble* rc = gin + static_cast<long long>(ya) * nx;
        double c0v = rc[lx0], c1v = rc[lx1];
        double ce = rc[lxe];
        const gdouble* rp = gin + static_cast<long long>((ya > 0 || tile) ? ya - 1 : ya) * nx;
        double p0v = rp[lx0], p1v = rp[lx1];
        if (tile) {  // halo rows pass through unchanged (step_task.hpp:78-84)
            if (ya == 0) {
                const long long o = -static_cast<long long>(nx);
                if (a0) g.out[o + x0] = g.in[o + x0];
                if (a1) g.out[o + x1] = g.in[o + x1];
            }
            if (yb == ny) {
                const long long o = static_cast<long long>(ny) * nx;
                if (a0) g.out[o + x0] = g.in[o + x0];
                if (a1) g.out[o + x1] = g.in[o + x1];
            }
        }

        // rows y+1 .. y+kAhead are in flight while row y is computed (the
        // values are only moved, never recomputed: the cell arithmetic is
        // unchanged); a row r is fetched when it is consumed (r <= yb) and
        // exists (r < ny, or the tile's halo row)
        const int r_last = (yb < ny || tile) ? yb : ny - 1;
        double q0[kAhead], q1[kAhead], qe[kAhead];
#pragma unroll
        for (int d = 0; d < kAhead; ++d) {
            const int r = ya + 1 + d;
            q0[d] = q1[d] = qe[d] = 0.0;
            if (r <= r_last) {
                const gdouble* rn = gin + static_cast<long long>(r) * nx;
                q0[d] = rn[lx0];
                q1[d] = rn[lx1];
                qe[d] = rn[lxe];
            }
        }
        // the laser of a plan-driven step, read after the rows are in
        // flight so that the two round trips overlap
        if (g.plan) {
            const LaserTick t =
                g.plan[(RED == RED_BATCH || RED == RED_ENSEMBLE) ? tick : *g.step];
            k.x_l = t.x_l;
            k.y_l = t.y_l;
            k.q_scale = t.q_scale;
        }
        const Beam bm = beam_of(k);
        for (int y = ya; y < yb; ++y) {
            const bool has_next = (y + 1 < ny || tile);
            const double n0v = has_next ? q0[0] : c0v, n1v = has_next ? q1[0] : c1v;
            const double ne = has_next ? qe[0] : ce;
#pragma unroll
            for (int d = 0; d + 1 < kAhead; ++d) {
                q0[d] = q0[d + 1];
                q1[d] = q1[d + 1];
                qe[d] = qe[d + 1];
            }
            const int r = y + 1 + kAhead;
            if (r <= r_last) {
                const gdouble* rn = gin + static_cast<long long>(r) * nx;
                q0[kAhead - 1] = rn[lx0];
                q1[kAhead - 1] = rn[lx1];
                qe[kAhead - 1] = rn[lxe];
            }
            const double up0 = lane_from_left(c0v), up1 = lane_from_left(c1v);
            const double dn0 = lane_from_right(c0v), dn1 = lane_from_right(c1v);
            const double last0 = lane_read(c0v, kWave - 1), first1 = lane_read(c1v, 0);
            const double xm[2] = {lane ? up0 : ce, lane ? up1 : last0};
            const double xp[2] = {lane < kWave - 1 ? dn0 : first1,
                                  lane < kWave - 1 ? dn1 : ce};
            const double cv[2] = {c0v, c1v}, pv[2] = {p0v, p1v}, nv[2] = {n0v, n1v};
            const int yg = g.y_off + y;
            const double ypos = static_cast<double>(yg) * k.dy;
            const bool far = beam_far(ypos, bm, k);
            gdouble* orow = gout + static_cast<long long>(y) * nx;
            bool mol[2] = {false, false};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int x = j ? x1 : x0;
                if (!(j ? a1 : a0)) continue;
                const double T = cv[j];
                const double xpos = static_cast<double>(x) * k.dx;
                double v, src = 0.0, lossv = 0.0;
                bool upd = true;
                if (g.bc == BC_DIRICHLET && (x == 0 || x == nx - 1 || y == 0 || y == ny - 1)) {
                    v = g.T_bc;
                    upd = false;
                } else if (PHYS == LPBF_PHYS_BASELINE) {
                    const int xe = (x == 0) ? -1 : (x == nx - 1 ? 1 : 0);
                    v = cell_base(T, xm[j], xp[j], pv[j], nv[j], xe, xpos, ypos, far, bm, k, rcp_k,
                                  src, lossv);
                } else {
                    const double Txm = (x == 0) ? T : xm[j];
                    const double Txp = (x == nx - 1) ? T : xp[j];
                    v = cell_ext(T, Txm, Txp, pv[j], nv[j], xpos, ypos, far, bm, k, rcp_k, src,
                                 lossv);
                }
                orow[x] = v;
                if (RED != RED_NONE) {
                    if (v > pa.tmax) {
                        pa.tmax = v;
                        bpos = 2 * y + j;
                    }
                    mol[j] = v > g.Tm;
                    if (RED == RED_STATS || RED == RED_RING || RED == RED_BATCH ||
                        RED == RED_ENSEMBLE) {
                        // held Dirichlet cells have src = lossv = 0 and add
                        // no temperature change
                        s_dT += v - (upd ? T : v);
                        s_src += src;
                        s_loss += lossv;
                    }
                }
            }
            if (RED != RED_NONE) {
                // molten cells and rows per wavefront from ballots: wave-
                // uniform, so they need no cross-lane reduction
                const unsigned long long b0 = __ballot(mol[0]), b1 = __ballot(mol[1]);
                molten += __popcll(b0) + __popcll(b1);
                if (b0 | b1) {
                    pa.ymin = min(pa.ymin, yg);
                    pa.ymax = max(pa.ymax, yg);
                }
            }
            p0v = c0v; p1v = c1v;
            c0v = n0v; c1v = n1v; ce = ne;
        }
    }

    if (RED == RED_NONE) return;
    if (bpos >= 0)
        pa.arg = static_cast<long long>(g.y_off + (bpos >> 1)) * nx + ((bpos & 1) ? xs1 : xs0);
    __shared__ PartialAcc s_pa[kThreads / kWave];
    __shared__ double s_sum[3][kThreads / kWave];
    __shared__ long long s_mol[kThreads / kWave];
    PartialAcc w = pa;   // ymin, ymax: already the wavefront's
    wave_hot(w.tmax, w.arg);
    double w_dT = 0.0, w_src = 0.0, w_loss = 0.0;
    const long long w_mol = molten;
    if (RED == RED_STATS || RED == RED_RING || RED == RED_BATCH || RED == RED_ENSEMBLE) {
        w_dT = wave_total(s_dT);
        w_src = wave_total(s_src);
        w_loss = wave_total(s_loss);
    }
    if (RED == RED_BATCH) {
        // one record per wavefront, written without a workgroup merge
        if (lane == 0)
            recs[bid * (kThreads / kWave) + wave] =
                StepStats{w.tmax, w.arg, w.ymin, w.ymax, w_mol, w_dT, w_src, w_loss};
        return;
    }
    if (lane == 0) {
        s_pa[wave] = w;
        s_sum[0][wave] = w_dT; s_sum[1][wave] = w_src; s_sum[2][wave] = w_loss;
        s_mol[wave] = w_mol;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PartialAcc b = s_pa[0];
        for (int i = 1; i < kThreads / kWave; ++i) b.merge(s_pa[i]);
        if (RED == RED_PARTIAL) {
            Partial* out = static_cast<Partial*>(red_out);
            out[blockIdx.x] = Partial{b.tmax, b.arg, b.ymin, b.ymax};
        } else {
            StepStats st;
            st.tmax = b.tmax; st.arg = b.arg; st.ymin = b.ymin; st.ymax = b.ymax;
            st.molten = s_mol[0]; st.sum_dT = s_sum[0][0];
            st.sum_src = s_sum[1][0]; st.sum_loss = s_sum[2][0];
            for (int i = 1; i < kThreads / kWave; ++i) {
                st.molten += s_mol[i];
                st.sum_dT += s_sum[0][i];
                st.sum_src += s_sum[1][i];
                st.sum_loss += s_sum[2][i];
            }
            recs[bid] = st;
        }
    }
    if (RED == RED_RING) {
        // last-block merge (no waiting: the workgroup holding the final ticket
        // does the step's remaining work)
        __shared__ int s_last, s_xhot;
        if (threadIdx.x == 0) {
            __threadfence();
            s_last = atomicAdd(g.ticket, 1) == static_cast<int>(gridDim.x) - 1;
        }
        __syncthreads();
        if (s_last) {
            __threadfence();
            const int s = *g.step;
            StepStats* slot = g.ring + s % g.ring_len;
            block_merge_stats(static_cast<const StepStats*>(red_out), gridDim.x, nx, g.ny_glob,
                              slot);
            if (threadIdx.x == 0) s_xhot = static_cast<int>(slot->arg % nx);
            __syncthreads();
            if (g.cols && g.cruise && g.cruise[s]) {
                double* col = g.cols + static_cast<long long>(s % g.ring_len) * ny;
                for (int y = threadIdx.x; y < ny; y += kThreads)
                    col[y] = g.out[static_cast<long long>(y) * nx + s_xhot];
            }
            if (threadIdx.x == 0) {
                *g.ticket = 0;
                *g.step = s + 1;
            }
        }
    }
    if (RED == RED_PARTIAL) {
        // records beyond this grid (the reference sizes `out` for 16x16
        // blocks) get the identity so a finalize over all of them is exact
        Partial* out = static_cast<Partial*>(red_out);
        const Partial id{-1.0, static_cast<long long>(nx) * g.ny_glob, g.ny_glob, -1};
        for (int i = gridDim.x + blockIdx.x * kThreads + threadIdx.x; i < red_fill;
             i += gridDim.x * kThreads)
            out[i] = id;
    }
}

// Reference reduce_kernel decomposition (kernels.cu:205-237): thread t of
// block b visits b*256+t, stepping by 256*nblocks; per-thread strict '>'.
__global__ __launch_bounds__(kThreads) void k_reduce_f64(const double* __restrict__ T,
                                                         int nx, int ny, double Tm,
                                                         Partial* __restrict__ out) {
    const long long N = static_cast<long long>(nx) * ny;
    PartialAcc p = PartialAcc::identity(nx, ny);
    const long long stride = static_cast<long long>(kThreads) * gridDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; i < N;
         i += stride) {
        const double v = T[i];
        if (v > p.tmax) { p.tmax = v; p.arg = i; }
        if (v > Tm) {
            const int yy = static_cast<int>(i / nx);
            if (yy < p.ymin) p.ymin = yy;
            if (yy > p.ymax) p.ymax = yy;
        }
    }
    __shared__ PartialAcc s_pa[kThreads / kWave];
    const PartialAcc w = wave_reduce(p);
    if ((threadIdx.x & (kWave - 1)) == 0) s_pa[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        PartialAcc b = s_pa[0];
        for (int i = 1; i < kThreads / kWave; ++i) b.merge(s_pa[i]);
        out[blockIdx.x] = Partial{b.tmax, b.arg, b.ymin, b.ymax};
    }
}

// One workgroup merges all partials: thread t folds t, t+256, ... in order,
// then a fixed tree.  The merge is exact, so this equals the serial loop.
__global__ __launch_bounds__(kThreads) void k_finalize(const Partial* __restrict__ parts,
                                                       int nblocks, int nx, int ny,
                                                       Partial* __restrict__ out) {
    PartialAcc p = PartialAcc::identity(nx, ny);
    for (int i = threadIdx.x; i < nblocks; i += kThreads) {
        const Partial q = parts[i];
        p.merge(PartialAcc{q.tmax, q.arg, q.ymin, q.ymax});
    }
    __shared__ PartialAcc s_pa[kThreads / kWave];
    const PartialAcc w = wave_reduce(p);
    if ((threadIdx.x & (kWave - 1)) == 0) s_pa[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        PartialAcc b = s_pa[0];
        for (int i = 1; i < kThreads / kWave; ++i) b.merge(s_pa[i]);
        out[0] = Partial{b.tmax, b.arg, b.ymin, b.ymax};
    }
}

// Finishes the last step of an ensemble batch (j = batch length) of every
// run, one workgroup per run.
__global__ __launch_bounds__(kThreads) void k_ensemble_flush(const EnsMember* __restrict__ runs,
                                                             const int* __restrict__ d_s0, int j,
                                                             int nblk, int nx, int ny) {
    const EnsMember& M = runs[blockIdx.x];
    const int s0 = *d_s0, s = s0 + j;
    if (j > 0 && s - 1 < M.n_steps)
        finish_batch_step(M.rec + ((s - 1) & 1) * nblk, nblk, nx, ny, j - 1, M.buf[s & 1],
                          M.cruise + s0, M.ring, M.cols, M.col_pitch);
}

// Same for StepStats; the FP64 sums use a fixed order (deterministic).
__global__ __launch_bounds__(kThreads) void k_batch_flush(const StepStats* __restrict__ rec,
                                                          int nblk, int nx, int ny, int j,
                                                          const double* __restrict__ field,
                                                          const unsigned char* __restrict__ cruise,
                                                          StepStats* __restrict__ ring,
                                                          double* __restrict__ cols) {
    finish_batch_step(rec, nblk, nx, ny, j, field, cruise, ring, cols, ny);
}

__global__ __launch_bounds__(kThreads) void k_stats_finalize(const StepStats* __restrict__ parts,
                                                             int nblocks, int nx, int ny,
                                                             StepStats* __restrict__ out) {
    block_merge_stats(parts, nblocks, nx, ny, out);
}

__global__ void k_extract_col(const double* __restrict__ T, int nx, int ny,
                              const Partial* __restrict__ glob, double* __restrict__ col) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    const int x_hot = static_cast<int>(glob->arg % nx);
    col[y] = T[static_cast<long long>(y) * nx + x_hot];
}

// The device exp exactly as the step kernels call it (same translation unit
// and flags), for measuring its distance to a host libm.
__global__ void k_probe_exp(const double* __restrict__ x, double* __restrict__ y, long long n) {
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        y[i] = exp(x[i]);
}

__global__ void k_noise_f64(double* __restrict__ T, long long n, unsigned long long seed,
                            double base, double amp) {
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        T[i] = base + amp * 2.0 * noise_u(seed, static_cast<unsigned long long>(i));
}

// Largest field the FP64 kernels accept (cells): 128 GiB of FP64, and every
// wavefront-tile count below stays far inside int range.
constexpr long long kMaxCells = 1LL << 34;

Region make_region(const double* in, double* out, int nx, int ny, int y_off, int bc,
                   double T_bc, double Tm, int ny_glob) {
    Region g;
    g.in = in; g.out = out; g.nx = nx; g.ny = ny; g.y_off = y_off; g.bc = bc;
    g.T_bc = T_bc; g.Tm = Tm; g.R = kRowsPerWave;
    g.nwx = (nx + kBlockCols - 1) / kBlockCols;
    g.ntasks = g.nwx * ((ny + g.R - 1) / g.R);
    g.ny_glob = ny_glob;
    g.plan = nullptr;
    g.step = nullptr;
    g.ticket = nullptr;
    g.ring = nullptr;
    g.ring_len = 1;
    g.cruise = nullptr;
    g.cols = nullptr;
    return g;
}

// Shape check shared by every launcher: the kernels trust nx and ny for
// every address they form, so out-of-range values are rejected here.
bool shape_ok(long long nx, long long ny) {
    return nx >= 2 && ny >= 1 && nx <= (1LL << 30) && ny <= (1LL << 30) && nx * ny <= kMaxCells;
}

int grid_of(const Region& g) {
    const int per = kThreads / kWave;
    return (g.ntasks + per - 1) / per;
}

template <int RED>
void launch_step(const Region& g, const StepConst& k, int phys, void* red, int fill,
                 hipStream_t s, int n_runs = 1) {
    // the batch forms have one more workgroup per step, which finishes the
    // previous step; the ensemble form repeats the grid for every run
    const int per = grid_of(g) + ((RED == RED_BATCH || RED == RED_ENSEMBLE) ? 1 : 0);
    const dim3 grid(per * n_runs), block(kThreads);
    Region gr = g;
    gr.rc = recips(k);
    if (phys == LPBF_PHYS_EXTENDED)
        hipLaunchKernelGGL((k_step_f64<LPBF_PHYS_EXTENDED, RED>), grid, block, 0, s, gr, k, red, fill);
    else
        hipLaunchKernelGGL((k_step_f64<LPBF_PHYS_BASELINE, RED>), grid, block, 0, s, gr, k, red, fill);
    check_launch();
}

bool valid_grid(const StepConst* k) { return k && shape_ok(k->nx, k->ny); }

// Band height of the plan-driven launch: halved from 32 rows until the grid
// has 4096 wavefront tiles (or 2 rows), so small plates still spread over
// the chip; a launch that is latency-bound gains more from parallel rows
// than it loses to the extra halo rows.
int plan_rows(int nx, int ny) {
    // shorter bands spread a small plate over more CUs and shorten each
    // wavefront's chain; down to one row (case0: 5.5 us per step at one
    // row against 6.3 at two)
    const int nwx = (nx + kBlockCols - 1) / kBlockCols;
    int R = kRowsPerWave;
    while (R > 1 && static_cast<long long>(nwx) * ((ny + R - 1) / R) < 4096) R /= 2;
    return R;
}

Region plan_region(const double* in, double* out, int nx, int ny, int bc, double T_bc,
                   double Tm) {
    Region g = make_region(in, out, nx, ny, 0, bc, T_bc, Tm, ny);
    g.R = plan_rows(nx, ny);
    g.ntasks = g.nwx * ((ny + g.R - 1) / g.R);
    return g;
}

// The ensemble launch: rows = 0 takes the single-run band height.
Region ensemble_region(int nx, int ny, int rows, int bc, double T_bc, double Tm) {
    Region g = plan_region(nullptr, nullptr, nx, ny, bc, T_bc, Tm);
    if (rows > 0) {
        g.R = rows;
        g.ntasks = g.nwx * ((ny + rows - 1) / rows);
    }
    return g;
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

extern "C" {

void lpbf_launch_step(const double* Tin, double* Tout, const StepConst* k, int phys,
                      hipStream_t stream) {
    if (!valid_grid(k) || !Tin || !Tout || (phys != 0 && phys != 1)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Region g = make_region(Tin, Tout, k->nx, k->ny, 0, BC_MIRROR, 0.0, 0.0, k->ny);
    launch_step<RED_NONE>(g, *k, phys, nullptr, 0, stream);
}

void lpbf_launch_step_reduce(const double* Tin, double* Tout, const StepConst* k, int phys,
                             double Tm, Partial* out, hipStream_t stream) {
    if (!valid_grid(k) || !Tin || !Tout || !out || (phys != 0 && phys != 1)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Region g = make_region(Tin, Tout, k->nx, k->ny, 0, BC_MIRROR, 0.0, Tm, k->ny);
    const int fill = ((k->nx + 15) / 16) * ((k->ny + 15) / 16);
    launch_step<RED_PARTIAL>(g, *k, phys, out, fill, stream);
}

void lpbf_launch_reduce(const double* T, int nx, int ny, double Tm, Partial* out,
                        int nblocks, hipStream_t stream) {
    if (!T || !out || nx < 1 || ny < 1 || static_cast<long long>(nx) * ny > kMaxCells ||
        nblocks < 1 || nblocks > (1 << 24)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_reduce_f64, dim3(nblocks), dim3(kThreads), 0, stream, T, nx, ny, Tm, out);
    check_launch();
}

void lpbf_launch_reduce_finalize(const Partial* parts, int nblocks, int nx, int ny,
                                 Partial* out, hipStream_t stream) {
    if (!parts || !out || nblocks < 0 || nx < 1 || ny < 1) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, stream, parts, nblocks, nx, ny, out);
    check_launch();
}

void lpbf_launch_extract_col(const double* T, int nx, int ny, const Partial* glob,
                             double* out_col, hipStream_t stream) {
    if (!T || !glob || !out_col || nx < 1 || ny < 1 ||
        static_cast<long long>(nx) * ny > kMaxCells) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const int bs = 128;
    hipLaunchKernelGGL(k_extract_col, dim3((ny + bs - 1) / bs), dim3(bs), 0, stream, T, nx, ny,
                       glob, out_col);
    check_launch();
}

void lpbf_launch_step_tile(const double* tile_in, double* tile_out, int owned_rows, int y0,
                           const StepConst* k, int phys, hipStream_t stream) {
    if (!k || !shape_ok(k->nx, static_cast<long long>(owned_rows) + 2) || owned_rows < 1 ||
        y0 < 0 || !tile_in || !tile_out || (phys != 0 && phys != 1)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Region g = make_region(tile_in + k->nx, tile_out + k->nx, k->nx, owned_rows, y0,
                                 BC_TILE, 0.0, 0.0, k->ny);
    launch_step<RED_NONE>(g, *k, phys, nullptr, 0, stream);
}

int lpbf_step_ex_nblocks(int nx, int ny) {
    if (!shape_ok(nx, ny)) return 0;
    const Region g = make_region(nullptr, nullptr, nx, ny, 0, BC_MIRROR, 0.0, 0.0, ny);
    return grid_of(g);
}

void lpbf_launch_step_ex(const double* Tin, double* Tout, const StepConst* k, int phys, int bc,
                         double T_bc, double Tm, StepStats* block_stats, hipStream_t stream) {
    if (!valid_grid(k) || !Tin || !Tout || (phys != 0 && phys != 1) ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Region g = make_region(Tin, Tout, k->nx, k->ny, 0, bc, T_bc, Tm, k->ny);
    if (block_stats)
        launch_step<RED_STATS>(g, *k, phys, block_stats, 0, stream);
    else
        launch_step<RED_NONE>(g, *k, phys, nullptr, 0, stream);
}

int lpbf_step_ex_plan_nblocks(int nx, int ny) {
    if (!shape_ok(nx, ny)) return 0;
    return grid_of(plan_region(nullptr, nullptr, nx, ny, BC_MIRROR, 0.0, 0.0));
}

void lpbf_launch_step_ex_plan(const double* Tin, double* Tout, const StepConst* k, int phys,
                              int bc, double T_bc, double Tm, const LaserTick* d_plan,
                              int* d_step, const unsigned char* d_cruise,
                              StepStats* block_stats, int* d_ticket, StepStats* ring,
                              int ring_len, double* cols, hipStream_t stream) {
    if (!valid_grid(k) || !Tin || !Tout || Tin == Tout || !d_plan || !d_step || !block_stats ||
        !d_ticket || !ring || ring_len < 1 || (phys != 0 && phys != 1) ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    Region g = plan_region(Tin, Tout, k->nx, k->ny, bc, T_bc, Tm);
    g.plan = d_plan;
    g.step = d_step;
    g.ticket = d_ticket;
    g.ring = ring;
    g.ring_len = ring_len;
    g.cruise = d_cruise;
    g.cols = cols;
    launch_step<RED_RING>(g, *k, phys, block_stats, 0, stream);
}

int lpbf_step_ex_batch_nblocks(int nx, int ny) {
    // one record per wavefront
    return lpbf_step_ex_plan_nblocks(nx, ny) * (kThreads / kWave);
}

void lpbf_launch_step_ex_batch(const double* Tin, double* Tout, const StepConst* k, int phys,
                               int bc, double T_bc, double Tm, const LaserTick* d_ticks, int j,
                               const unsigned char* d_cruise, StepStats* block_stats,
                               StepStats* ring, double* cols, hipStream_t stream) {
    if (!valid_grid(k) || !Tin || !Tout || Tin == Tout || !d_ticks || j < 0 || !block_stats ||
        !ring || (phys != 0 && phys != 1) ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    Region g = plan_region(Tin, Tout, k->nx, k->ny, bc, T_bc, Tm);
    g.plan = d_ticks;
    g.ring = ring;
    g.ring_len = 0;
    g.cruise = d_cruise;
    g.cols = cols;
    launch_step<RED_BATCH>(g, *k, phys, block_stats, j, stream);
}

void lpbf_launch_step_ex_batch_flush(const double* T, int nx, int ny, int j,
                                     const unsigned char* d_cruise, const StepStats* block_stats,
                                     StepStats* ring, double* cols, hipStream_t stream) {
    if (!T || !shape_ok(nx, ny) || j < 1 || !block_stats || !ring) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const int nblk = lpbf_step_ex_batch_nblocks(nx, ny);
    hipLaunchKernelGGL(k_batch_flush, dim3(1), dim3(kThreads), 0, stream,
                       block_stats + ((j - 1) & 1) * nblk, nblk, nx, ny, j - 1, T, d_cruise, ring,
                       cols);
    check_launch();
}

int lpbf_step_ensemble_rows(int nx, int ny, int n_runs) {
    if (!shape_ok(nx, ny) || n_runs < 1) return 0;
    // measured on 240x120 rows (1, 64 and 256 of them): short bands pay the
    // per-wavefront set-up and reductions, few wavefronts leave SIMDs idle.
    // Few rows (under ~1k wavefronts at 4-row bands): the shortest bands
    // that keep ~4k wavefronts; otherwise 4-row bands up to ~8k wavefronts
    // in all, longer ones beyond
    const long long nwx = (nx + kBlockCols - 1) / kBlockCols;
    auto waves = [&](int R) { return nwx * ((ny + R - 1) / R) * n_runs; };
    int R = 4;
    if (waves(4) < 1024) {
        R = 1;
        while (R < 4 && waves(R) > 4096) R *= 2;
        return R;
    }
    while (R < 32 && waves(R) > 8192) R *= 2;
    return R;
}

int lpbf_step_ensemble_nblocks(int nx, int ny, int rows) {
    if (!shape_ok(nx, ny) || rows < 0 || rows > kMaxEnsembleRows) return 0;
    return grid_of(ensemble_region(nx, ny, rows, BC_MIRROR, 0.0, 0.0));
}

void lpbf_launch_step_ensemble(const lpbf_ensemble_run* d_runs, int n_runs, int nx, int ny,
                               int rows, int phys, int bc, double T_bc, double Tm,
                               const int* d_s0, int j, hipStream_t stream) {
    const int nblk = lpbf_step_ensemble_nblocks(nx, ny, rows);
    if (!d_runs || n_runs < 1 || n_runs > 65535 || nblk <= 0 || !d_s0 || j < 0 ||
        (phys != 0 && phys != 1) || (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET) ||
        static_cast<long long>(nblk + 1) * n_runs > (1LL << 31) - 1) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    Region g = ensemble_region(nx, ny, rows, bc, T_bc, Tm);
    g.step = const_cast<int*>(d_s0);
    StepConst k{};
    launch_step<RED_ENSEMBLE>(g, k, phys, const_cast<lpbf_ensemble_run*>(d_runs), j, stream,
                              n_runs);
}

void lpbf_launch_step_ensemble_flush(const lpbf_ensemble_run* d_runs, int n_runs, int nx, int ny,
                                     int rows, const int* d_s0, int j, hipStream_t stream) {
    const int nblk = lpbf_step_ensemble_nblocks(nx, ny, rows);
    if (!d_runs || n_runs < 1 || n_runs > 65535 || nblk <= 0 || !d_s0 || j < 1) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_ensemble_flush, dim3(n_runs), dim3(kThreads), 0, stream, d_runs, d_s0,
                       j, nblk, nx, ny);
    check_launch();
}

void lpbf_launch_stats_finalize(const StepStats* parts, int nblocks, int nx, int ny,
                                StepStats* out, hipStream_t stream) {
    if (!parts || !out || nblocks < 0 || nx < 1 || ny < 1) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_stats_finalize, dim3(1), dim3(kThreads), 0, stream, parts, nblocks, nx,
                       ny, out);
    check_launch();
}

void lpbf_launch_noise_f64(double* T, long long n, unsigned long long seed, double base,
                           double amp, hipStream_t stream) {
    if (!T || n < 0) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    if (n == 0) return;
    const long long want = (n + kThreads - 1) / kThreads;
    const int grid = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_noise_f64, dim3(grid), dim3(kThreads), 0, stream, T, n, seed, base, amp);
    check_launch();
}

void lpbf_probe_exp_f64(const double* x, double* y, long long n, hipStream_t stream) {
    if (!x || !y || n < 0) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    if (n == 0) return;
    const long long want = (n + kThreads - 1) / kThreads;
    const int grid = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_probe_exp, dim3(grid), dim3(kThreads), 0, stream, x, y, n);
    check_launch();
}

}  // extern "C"
