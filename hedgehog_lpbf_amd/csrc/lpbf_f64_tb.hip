// SPDX-License-Identifier: 0BSD
// FP64 steps with temporal blocking: one HBM pass advances the field by
// `depth` (1..4) steps.  Every cell of every step evaluates the same
// expression as the single-step FP64 kernels (lpbf_f64_cell.hpp), on the
// same operands, so the field after a launch is bitwise the field of `depth`
// lpbf_launch_step_ex calls; only the order of the per-step FP64 energy sums
// differs.
//
// Schedule (MI355X-first):
//   * a wavefront owns a strip of 128 loaded columns, two adjacent columns
//     per lane (one 16-byte access per lane, 1 KiB per row); strips overlap
//     by ceil(depth/2) lanes on each side, because every level invalidates
//     one more column at each strip edge;
//   * rows stream top to bottom through a register pipeline: loading row i
//     advances level j to row i-j, so `depth` steps cost one read and one
//     write of the field plus the 2*depth recomputed band-halo rows;
//   * level 0 holds a six-row ring (rows i-2 .. i+3, three rows in flight),
//     every other level a three-row ring, all indexed by a compile-time
//     phase; the row loop runs six phases per trip so no register moves;
//   * x neighbours move between lanes by DPP wave shifts, y neighbours come
//     from the rings; the first/last domain rows use the ghost = centre rule
//     of the reference (step_task.hpp:87-92), the first/last domain columns
//     its one-sided forms (cell_base);
//   * per-step statistics (max with index, molten rows, energy sums) are
//     reduced per wavefront and workgroup and merged by a second kernel.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lpbf_common.hpp"
#include "lpbf_f64_cell.hpp"
#include "lpbf_kernels.h"

namespace lpbf_amd {
namespace {

constexpr int kTbThreads = 256;               // 4 wavefronts, one strip-band each
constexpr int kTbWaves = kTbThreads / kWave;
constexpr int kTbCols = 2 * kWave;            // loaded columns per strip
constexpr int kTbMaxDepth = 4;

template <int K>
struct TbShape {
    static constexpr int ov = (K + 1) / 2;            // overlap lanes per side
    static constexpr int out = kTbCols - 4 * ov;      // columns written per strip
};

struct TbArgs {
    const double* in;
    double* out;
    int nx, ny;
    int nstrips, R, ntasks, nwg;
    double T_bc, Tm;
    StepStats* part;                                  // [depth][nwg] or null
};

struct D2 {
    double x, y;
};

// EVEN = true promises an even nx with mirror edges: lane columns start at
// even positions, so a lane's two columns are both inside the field or both
// outside it, and with mirror edges every owned cell counts in the sums.
// Ownership is then one flag per lane, applied once before the reductions;
// the per-cell statistics run unpredicated, the running maximum takes the
// row pair's larger value (the first column on ties, as the strict > of the
// cell-by-cell order) and the molten cells and rows are counted for the
// wavefront from ballots.
template <int K, int PHYS, int BC, bool EVEN>
__global__ __launch_bounds__(kTbThreads) __attribute__((amdgpu_waves_per_eu(K <= 3 ? 3 : 2))) void k_step_f64_tb(TbArgs a, StepConst k,
                                                            const LaserTick* __restrict__ plan,
                                                            int step0) {
    using Sh = TbShape<K>;
    constexpr bool kLaneOwned = EVEN && BC != LPBF_BC_DIRICHLET;
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform: the band, its rows and the row tests live in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int wg = xcd_remap(blockIdx.x, a.nwg);
    const int task = wg * kTbWaves + wave;
    const int nx = a.nx, ny = a.ny;

    Beam bm[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const LaserTick t = plan[step0 + j];
        bm[j] = Beam{t.x_l, t.y_l, t.q_scale};
    }
    const Recips rc = recips(k);

    // Per-lane statistics of each step.  A lane meets its cells in
    // increasing row-major order, so a strict > keeps the smallest index
    // among equal maxima; the index is kept as 2*row + (second column).
    double best[K];
    int bpos[K], ymin[K], ymax[K], molten[K];
    double s_dT[K], s_src[K], s_loss[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        best[j] = -1.0;
        bpos[j] = -1;
        ymin[j] = ny;
        ymax[j] = -1;
        molten[j] = 0;
        s_dT[j] = s_src[j] = s_loss[j] = 0.0;
    }
    int xcol[2] = {0, 0};     // the lane's two columns (for the index)
    bool lane_owns = false;   // kLaneOwned: the lane's cells count
    unsigned long long own_mask = 0;   // kLaneOwned: the owning lanes (wave-uniform)

    if (task < a.ntasks) {
        const int band = task / a.nstrips, strip = task % a.nstrips;
        const int L0 = strip * Sh::out - 2 * Sh::ov;
        const int x0 = L0 + 2 * lane, x1 = x0 + 1;
        const bool in0 = x0 >= 0 && x0 < nx, in1 = x1 >= 0 && x1 < nx;
        const bool own_lane = lane >= Sh::ov && lane < kWave - Sh::ov;
        const bool own0 = own_lane && in0, own1 = own_lane && in1;
        lane_owns = own0;
        own_mask = __builtin_amdgcn_ballot_w64(own0);
        const int ya = band * a.R;
        const int yb = min(ya + a.R, ny);
        const int y_lo = max(ya - K, 0);
        const int y_hi = min(yb + K, ny);
        const int i_end = yb - 1 + K;     // level K row yb-1 is produced at i = yb-1+K
        xcol[0] = x0;
        xcol[1] = x1;
        const double xpos0 = static_cast<double>(x0) * k.dx;
        const double xpos1 = static_cast<double>(x1) * k.dx;
        const int xe0 = (x0 == 0) ? -1 : (x0 == nx - 1 ? 1 : 0);
        const int xe1 = (x1 == 0) ? -1 : (x1 == nx - 1 ? 1 : 0);

        // rows past y_hi are never consumed: the last row is loaded instead
        auto load_row = [&](int i) {
            const double* p = a.in + static_cast<long long>(min(i, y_hi - 1)) * nx;
            D2 v;
            v.x = in0 ? p[x0] : 0.0;
            v.y = in1 ? p[x1] : 0.0;
            return v;
        };

        D2 Z[6];
        D2 W[K > 1 ? K - 1 : 1][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) Z[q] = load_row(y_lo + q);

        auto win_new = [&](auto PHC, auto JC) -> D2& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[ph % 6];
            else return W[j - 1][ph % 3];
        };
        auto win_mid = [&](auto PHC, auto JC) -> D2& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[(ph + 5) % 6];
            else return W[j - 1][(ph + 2) % 3];
        };
        auto win_old = [&](auto PHC, auto JC) -> D2& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[(ph + 4) % 6];
            else return W[j - 1][(ph + 1) % 3];
        };

        // one cell of step j (0-based within the launch); returns the new
        // value and adds its energy terms to that step's sums
        auto cell = [&](auto SC, double T, double Txm, double Txp, double Tym, double Typ, int x,
                        int xe, double xpos, double ypos, bool far, int r, bool own) {
            constexpr int j = decltype(SC)::value;
            double v, src = 0.0, lossv = 0.0;
            bool upd = true;
            if (BC == LPBF_BC_DIRICHLET && (x == 0 || x == nx - 1 || r == 0 || r == ny - 1)) {
                v = a.T_bc;
                upd = false;
            } else if (PHYS == LPBF_PHYS_BASELINE) {
                v = cell_base(T, Txm, Txp, Tym, Typ, xe, xpos, ypos, far, bm[j], k, rc, src,
                              lossv);
            } else {
                const double Txm_g = (x == 0) ? T : Txm;        // ghost = centre
                const double Txp_g = (x == nx - 1) ? T : Txp;
                v = cell_ext(T, Txm_g, Txp_g, Tym, Typ, xpos, ypos, far, bm[j], k, rc, src,
                             lossv);
            }
            if (kLaneOwned ? own : (own && upd)) {   // kLaneOwned: own is the row test
                s_dT[j] += v - T;
                s_src[j] += src;
                s_loss[j] += lossv;
            }
            return v;
        };

        // level j (1..K) at loop row i: row r = i - j from level j-1's window
        auto level = [&](auto PHC, auto JC, int i) {
            constexpr int j = decltype(JC)::value;
            using JP = std::integral_constant<int, j - 1>;
            const int r = i - j;
            // level j is needed on rows [lo, hi): the band and the halo the
            // remaining K - j levels consume (wave-uniform test)
            const int lo = max(ya - (K - j), 0);
            const int hi = min(yb + (K - j), ny);
            if (r < lo || r >= hi) return;
            const D2 c = win_mid(PHC, JP{});
            const D2 up = (BC != LPBF_BC_DIRICHLET && r == 0) ? c : win_old(PHC, JP{});
            const D2 dn = (BC != LPBF_BC_DIRICHLET && r == ny - 1) ? c : win_new(PHC, JP{});
            const double xl = lane_from_left(c.y);    // column x0 - 1
            const double xr = lane_from_right(c.x);   // column x1 + 1
            const double ypos = static_cast<double>(r) * k.dy;
            const bool far = beam_far(ypos, bm[j - 1], k);
            const bool owned_row = r >= ya && r < yb;
            D2 res;
            res.x = cell(JP{}, c.x, xl, c.y, up.x, dn.x, x0, xe0, xpos0, ypos, far, r,
                         kLaneOwned ? owned_row : (owned_row && own0));
            res.y = cell(JP{}, c.y, c.x, xr, up.y, dn.y, x1, xe1, xpos1, ypos, far, r,
                         kLaneOwned ? owned_row : (owned_row && own1));
            if (kLaneOwned && owned_row) {
                constexpr int s = j - 1;
                const double m = __builtin_fmax(res.x, res.y);
                if (m > best[s]) {
                    best[s] = m;
                    bpos[s] = (res.x == m) ? 2 * r : 2 * r + 1;
                }
                // molten cells and rows of the owning lanes, counted for the
                // wavefront from ballots (scalar registers)
                if (__builtin_amdgcn_ballot_w64(m > a.Tm) & own_mask) {
                    molten[s] +=
                        __builtin_popcountll(__builtin_amdgcn_ballot_w64(res.x > a.Tm) & own_mask) +
                        __builtin_popcountll(__builtin_amdgcn_ballot_w64(res.y > a.Tm) & own_mask);
                    ymin[s] = min(ymin[s], r);
                    ymax[s] = r;
                }
            } else if (owned_row) {
                constexpr int s = j - 1;
                const bool o0 = own0 && res.x > best[s];
                best[s] = o0 ? res.x : best[s];
                bpos[s] = o0 ? 2 * r : bpos[s];
                const bool o1 = own1 && res.y > best[s];
                best[s] = o1 ? res.y : best[s];
                bpos[s] = o1 ? 2 * r + 1 : bpos[s];
                const int m = (own0 && res.x > a.Tm) + (own1 && res.y > a.Tm);
                molten[s] += m;
                if (m) {
                    ymin[s] = min(ymin[s], r);
                    ymax[s] = r;
                }
            }
            if constexpr (j == K) {
                if (owned_row) {
                    double* o = a.out + static_cast<long long>(r) * nx;
                    if (own0) o[x0] = res.x;
                    if (own1) o[x1] = res.y;
                }
            } else {
                win_new(PHC, JC) = res;
            }
        };

        auto body = [&](auto PHC, int i) {
            constexpr int ph = decltype(PHC)::value;
            Z[(ph + 3) % 6] = load_row(i + 3);
            level(PHC, std::integral_constant<int, 1>{}, i);
            if constexpr (K >= 2) level(PHC, std::integral_constant<int, 2>{}, i);
            if constexpr (K >= 3) level(PHC, std::integral_constant<int, 3>{}, i);
            if constexpr (K >= 4) level(PHC, std::integral_constant<int, 4>{}, i);
        };

        // six rows per trip, one per ring phase; rows past i_end fail every
        // level's row test and load clamped rows.  The wavefront lowers its
        // issue priority at the quarter marks of its rows, so the waves that
        // are behind take the issue slots and a SIMD's waves finish together
        const int span = i_end + 1 - y_lo;
        const int q1 = y_lo + span / 4, q2 = y_lo + span / 2, q3 = y_lo + (3 * span) / 4;
        for (int i = y_lo; i <= i_end; i += 6) {
            if (i >= q3) __builtin_amdgcn_s_setprio(0);
            else if (i >= q2) __builtin_amdgcn_s_setprio(1);
            else if (i >= q1) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(3);
            body(std::integral_constant<int, 0>{}, i);
            body(std::integral_constant<int, 1>{}, i + 1);
            body(std::integral_constant<int, 2>{}, i + 2);
            body(std::integral_constant<int, 3>{}, i + 3);
            body(std::integral_constant<int, 4>{}, i + 4);
            body(std::integral_constant<int, 5>{}, i + 5);
        }
    }

    if (!a.part) return;
    if (kLaneOwned && !lane_owns) {
#pragma unroll
        for (int j = 0; j < K; ++j) {   // molten and the row extent are per wavefront
            bpos[j] = -1;
            s_dT[j] = s_src[j] = s_loss[j] = 0.0;
        }
    }
    // per level: wavefront then workgroup merge -> a.part[j*nwg + wg]
    __shared__ PartialAcc s_pa[kTbWaves];
    __shared__ double s_sum[3][kTbWaves];
    __shared__ long long s_m[kTbWaves];
    const long long N = static_cast<long long>(nx) * ny;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        PartialAcc p;
        p.tmax = bpos[j] < 0 ? -1.0 : best[j];
        p.arg = bpos[j] < 0 ? N
                            : static_cast<long long>(bpos[j] >> 1) * nx + xcol[bpos[j] & 1];
        p.ymin = ymin[j];
        p.ymax = ymax[j];
        const PartialAcc w = wave_reduce(p);
        const double wd = wave_sum(s_dT[j]), ws = wave_sum(s_src[j]), wl = wave_sum(s_loss[j]);
        const long long wm = kLaneOwned ? static_cast<long long>(molten[j])
                                        : wave_sum(static_cast<long long>(molten[j]));
        if (lane == 0) {
            s_pa[wave] = w;
            s_sum[0][wave] = wd;
            s_sum[1][wave] = ws;
            s_sum[2][wave] = wl;
            s_m[wave] = wm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            PartialAcc b = s_pa[0];
            StepStats st;
            st.molten = s_m[0]; st.sum_dT = s_sum[0][0]; st.sum_src = s_sum[1][0];
            st.sum_loss = s_sum[2][0];
            for (int q = 1; q < kTbWaves; ++q) {
                b.merge(s_pa[q]);
                st.molten += s_m[q]; st.sum_dT += s_sum[0][q];
                st.sum_src += s_sum[1][q]; st.sum_loss += s_sum[2][q];
            }
            st.tmax = b.tmax; st.arg = b.arg; st.ymin = b.ymin; st.ymax = b.ymax;
            a.part[static_cast<long long>(j) * a.nwg + wg] = st;
        }
        __syncthreads();
    }
}

// One workgroup per level merges that level's workgroup records.
__global__ __launch_bounds__(256) void k_tb_stats_finalize(const StepStats* __restrict__ part,
                                                           int nwg, int nx, int ny,
                                                           StepStats* __restrict__ out) {
    block_merge_stats(part + static_cast<long long>(blockIdx.x) * nwg, nwg, nx, ny,
                      out + blockIdx.x);
}

// Built-in band height: the grid keeps about four rounds of resident
// wavefronts (1024 SIMDs, 3 waves each), never below 8*depth rows, where
// the 2*depth recomputed halo rows would dominate.
int tb_rows(int nx, int ny, int depth) {
    const long long out = kTbCols - 4 * ((depth + 1) / 2);
    const long long nstrips = (nx + out - 1) / out;
    const long long bands = (4 * 3 * 1024 + nstrips - 1) / nstrips;
    int R = static_cast<int>((ny + bands - 1) / bands);
    const int lo = 8 * depth;
    if (R < lo) R = lo;
    if (R > ny) R = ny;
    return R < 1 ? 1 : R;
}

struct TbGeom {
    int nstrips, R, ntasks, nwg;
};

TbGeom tb_geometry(int nx, int ny, int depth, int rows) {
    TbGeom g;
    const int out = kTbCols - 4 * ((depth + 1) / 2);
    g.nstrips = (nx + out - 1) / out;
    g.R = rows > 0 ? rows : tb_rows(nx, ny, depth);
    g.ntasks = g.nstrips * ((ny + g.R - 1) / g.R);
    g.nwg = (g.ntasks + kTbWaves - 1) / kTbWaves;
    return g;
}

template <int K, int PHYS>
void launch_tb_k(int bc, const TbArgs& a, const StepConst& k, const LaserTick* plan, int step0,
                 hipStream_t s) {
    if (bc == LPBF_BC_DIRICHLET)
        hipLaunchKernelGGL((k_step_f64_tb<K, PHYS, LPBF_BC_DIRICHLET, false>), dim3(a.nwg),
                           dim3(kTbThreads), 0, s, a, k, plan, step0);
    else if (a.nx % 2 == 0)
        hipLaunchKernelGGL((k_step_f64_tb<K, PHYS, LPBF_BC_MIRROR, true>), dim3(a.nwg),
                           dim3(kTbThreads), 0, s, a, k, plan, step0);
    else
        hipLaunchKernelGGL((k_step_f64_tb<K, PHYS, LPBF_BC_MIRROR, false>), dim3(a.nwg),
                           dim3(kTbThreads), 0, s, a, k, plan, step0);
}

template <int PHYS>
void launch_tb_p(int depth, int bc, const TbArgs& a, const StepConst& k, const LaserTick* plan,
                 int step0, hipStream_t s) {
    switch (depth) {
        case 1: launch_tb_k<1, PHYS>(bc, a, k, plan, step0, s); break;
        case 2: launch_tb_k<2, PHYS>(bc, a, k, plan, step0, s); break;
        case 3: launch_tb_k<3, PHYS>(bc, a, k, plan, step0, s); break;
        default: launch_tb_k<4, PHYS>(bc, a, k, plan, step0, s); break;
    }
}

bool tb_shape_ok(long long nx, long long ny) {
    return nx >= 2 && ny >= 1 && nx <= (1LL << 30) && ny <= (1LL << 30) &&
           nx * ny <= (1LL << 34);
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

extern "C" {

size_t lpbf_f64_blocked_workspace_bytes(int nx, int ny, int depth, int rows) {
    if (!tb_shape_ok(nx, ny) || depth < 1 || depth > kTbMaxDepth || rows < 0) return 0;
    // every launch of depth d <= depth writes d * nwg(d) records, and with the
    // built-in band height (rows == 0) a shorter depth can mean more
    // workgroups, since the minimum band height grows with the depth
    size_t need = 0;
    for (int d = 1; d <= depth; ++d) {
        const TbGeom g = tb_geometry(nx, ny, d, rows);
        const size_t b = sizeof(StepStats) * static_cast<size_t>(d) * static_cast<size_t>(g.nwg);
        if (b > need) need = b;
    }
    return need;
}

void lpbf_launch_step_f64_blocked(const double* Tin, double* Tout, const StepConst* k, int phys,
                                  int bc, double T_bc, double Tm, const LaserTick* d_plan,
                                  int step0, int depth, int rows, StepStats* stats, void* work,
                                  hipStream_t stream) {
    if (!k || !tb_shape_ok(k->nx, k->ny) || !Tin || !Tout || Tin == Tout || !d_plan ||
        step0 < 0 || depth < 1 || depth > kTbMaxDepth || rows < 0 ||
        (phys != LPBF_PHYS_BASELINE && phys != LPBF_PHYS_EXTENDED) ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET) || (stats && !work)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const TbGeom g = tb_geometry(k->nx, k->ny, depth, rows);
    TbArgs a;
    a.in = Tin; a.out = Tout; a.nx = k->nx; a.ny = k->ny;
    a.nstrips = g.nstrips; a.R = g.R; a.ntasks = g.ntasks; a.nwg = g.nwg;
    a.T_bc = T_bc; a.Tm = Tm;
    a.part = stats ? static_cast<StepStats*>(work) : nullptr;
    if (phys == LPBF_PHYS_EXTENDED)
        launch_tb_p<LPBF_PHYS_EXTENDED>(depth, bc, a, *k, d_plan, step0, stream);
    else
        launch_tb_p<LPBF_PHYS_BASELINE>(depth, bc, a, *k, d_plan, step0, stream);
    check_launch();
    if (stats) {
        hipLaunchKernelGGL(k_tb_stats_finalize, dim3(depth), dim3(256), 0, stream, a.part, a.nwg,
                           a.nx, a.ny, stats);
        check_launch();
    }
}

int lpbf_f64_blocked_autotune(const StepConst* k, int phys, int bc, double T_bc, double Tm,
                              const LaserTick* d_plan, int n_plan, int max_depth,
                              hipStream_t stream, int* best_depth, int* best_rows,
                              double* best_ms_per_step) {
    if (!k || !tb_shape_ok(k->nx, k->ny) || !d_plan || max_depth < 1 ||
        max_depth > kTbMaxDepth || n_plan < max_depth || !best_depth || !best_rows ||
        (phys != LPBF_PHYS_BASELINE && phys != LPBF_PHYS_EXTENDED) ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return LPBF_ERR_INVALID_ARGUMENT;
    }
    const int nx = k->nx, ny = k->ny;
    // three resident waves per SIMD up to depth 3, 1024 SIMDs
    auto rows_for = [&](int depth) {
        const int out = kTbCols - 4 * ((depth + 1) / 2);
        return band_candidates(ny, (nx + out - 1) / out, 3 * 1024, {0, 16, 32, 48, 64, 96},
                               2 * depth);
    };
    size_t work_bytes = 0;
    for (int d = 1; d <= max_depth; ++d)
        for (int R : rows_for(d)) {
            const size_t b = lpbf_f64_blocked_workspace_bytes(nx, ny, d, R);
            if (b > work_bytes) work_bytes = b;
        }
    double* a = nullptr;
    double* b = nullptr;
    void* work = nullptr;
    StepStats* st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t field = sizeof(double) * static_cast<size_t>(nx) * ny;
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&a), field);
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&b), field);
    if (err == hipSuccess) err = hipMalloc(&work, work_bytes);
    if (err == hipSuccess)
        err = hipMalloc(reinterpret_cast<void**>(&st), sizeof(StepStats) * kTbMaxDepth);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    double best = 1e300;
    int bd = 1, br = 0;
    if (err == hipSuccess) {
        lpbf_launch_noise_f64(a, static_cast<long long>(nx) * ny, 1, 300.0, 0.5, stream);
        const int reps = 10;   // long enough that launch-to-launch noise does not pick the tiling
        for (int depth = 1; depth <= max_depth && err == hipSuccess; ++depth) {
            for (int R : rows_for(depth)) {
                lpbf_launch_step_f64_blocked(a, b, k, phys, bc, T_bc, Tm, d_plan, 0, depth, R,
                                             st, work, stream);   // warm-up
                err = hipEventRecord(e0, stream);
                if (err != hipSuccess) break;
                for (int q = 0; q < reps; ++q)
                    lpbf_launch_step_f64_blocked((q & 1) ? a : b, (q & 1) ? b : a, k, phys, bc,
                                                 T_bc, Tm, d_plan, 0, depth, R, st, work,
                                                 stream);
                err = hipEventRecord(e1, stream);
                if (err == hipSuccess) err = hipEventSynchronize(e1);
                float ms = 0.f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                if (err != hipSuccess) break;
                const double per_step = ms / (reps * depth);
                if (per_step < best) { best = per_step; bd = depth; br = R; }
            }
        }
    }
    // release everything; the first error (if any) is the one reported
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (work) (void)hipFree(work);
    if (st) (void)hipFree(st);
    if (err != hipSuccess) {
        record_error(static_cast<int>(err));
        return static_cast<int>(err);
    }
    *best_depth = bd;
    *best_rows = br;
    if (best_ms_per_step) *best_ms_per_step = best;
    return 0;
}

}  // extern "C"
