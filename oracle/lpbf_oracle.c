/* SPDX-License-Identifier: 0BSD
 *
 * CPU oracle for the hedgehog_lpbf z-averaged 2D FD heat step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the tests, the
 * smoke check and the benchmark's CPU-baseline leg compare the HIP path
 * against; nothing in the product links or calls it.
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   ora_dt_cfl / ora_make_const     case0 dt_s and n_steps (meta.json:11-12);
 *                                   bit-equal to the reference headers when
 *                                   oracle/_ref builds
 *   ora_step / ora_step_tile        case0 metrics.csv rows + meta.json peaks,
 *                                   first-step closed form, eigenmode decay,
 *                                   tiling independence, numpy brute force
 *   ora_measure / ora_reduce_*      case0 W_um column, brute force, block-size
 *                                   independence
 *   ora_build_plan                  case0 x_laser_um column (incl. the wrap)
 *   ora_k_of / ora_cp_of / ...      table knots, linear segments, kmult = 1
 *                                   identity; bit-equal to the reference
 *                                   headers when oracle/_ref builds
 *   ora_step_ex (Dirichlet, stats)  sine-mode decay, energy identity,
 *                                   reduction to ora_step for BC mirror
 *   ora_noise                       SplitMix64 published first output
 *   ora_raster_plan                 closed-form positions at segment ends
 *   ora_step_ex (extended stats)    source sum = change of sum(T'-T) with the
 *                                   beam on, linear in q_scale; loss-model
 *                                   differences; zero loss at equilibrium
 *   ora_step_rows_f32               (the FP32 CPU baseline, no parity role)
 *                                   within FP32 rounding of ora_step on a
 *                                   field whose every term is far above it
 */
#include "lpbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORA_PI 3.14159265358979323846          /* M_PI of <cmath> */
#define ORA_SIGMA_SB 5.670374419e-8            /* params.hpp:11 */

/* ---- constants and dt (params.hpp:13-47) ------------------------------- */

void ora_defaults(OraMaterial* m, OraProcess* p, OraDomain* d) {
    if (m) {  /* params.hpp:13-21 */
        m->k = 25.0; m->rho = 8190.0; m->cp = 565.0; m->Tm = 1609.0;
        m->eps = 0.40; m->eta = 0.35;
    }
    if (p) {  /* params.hpp:23-33 */
        p->P = 285.0; p->V = 0.96; p->r0 = 33.5e-6; p->h = 40.0e-6;
        p->hConv = 100.0; p->dSub = 30.0e-6; p->Tamb = 298.15; p->Tsub = 298.15;
        p->radiation = true;
    }
    if (d) {  /* params.hpp:35-40 */
        d->Lx = 1.6e-3; d->Ly = 0.6e-3; d->nx = 240; d->ny = 120;
    }
}

/* AMB2022-03 case table (main.cpp:97-109).  Returns 0, or -1 if unknown. */
int ora_apply_case(const char* id, OraProcess* p) {
    if      (!strcmp(id, "0"))   { p->P = 285.0; p->V = 0.96; p->r0 = 33.5e-6; }
    else if (!strcmp(id, "1.1")) { p->P = 285.0; p->V = 0.96; p->r0 = 24.5e-6; }
    else if (!strcmp(id, "1.2")) { p->P = 285.0; p->V = 0.96; p->r0 = 41.0e-6; }
    else if (!strcmp(id, "2.1")) { p->P = 285.0; p->V = 1.20; p->r0 = 33.5e-6; }
    else if (!strcmp(id, "2.2")) { p->P = 285.0; p->V = 0.80; p->r0 = 33.5e-6; }
    else if (!strcmp(id, "3.1")) { p->P = 325.0; p->V = 0.96; p->r0 = 33.5e-6; }
    else if (!strcmp(id, "3.2")) { p->P = 245.0; p->V = 0.96; p->r0 = 33.5e-6; }
    else return -1;
    return 0;
}

double ora_alpha(const OraMaterial* m) { return m->k / (m->rho * m->cp); }  /* params.hpp:20 */
double ora_dx(const OraDomain* d) { return d->Lx / (d->nx - 1); }           /* params.hpp:38 */
double ora_dy(const OraDomain* d) { return d->Ly / (d->ny - 1); }           /* params.hpp:39 */

/* params.hpp:44-47 */
double ora_dt_cfl(const OraMaterial* m, const OraDomain* d) {
    const double a = ora_alpha(m);
    const double dx = ora_dx(d), dy = ora_dy(d);
    return 0.35 / (a * (1.0 / (dx * dx) + 1.0 / (dy * dy)));
}

/* ---- IN718 extended physics (params.hpp:61-132) ------------------------- */

static const double T_SOL = 1533.0, T_LIQ = 1609.0, LF = 210.0e3;  /* :63-65 */
#define ORA_NPTS 9
static const double T_PTS[ORA_NPTS]  = {298, 473, 673, 873, 1073,
                                        1273, 1473, 1533, 1609};   /* :68-69 */
static const double K_PTS[ORA_NPTS]  = {8.9, 12.9, 15.8, 18.7, 21.9,
                                        25.6, 28.6, 29.3, 29.6};   /* :70-71 */
static const double CP_PTS[ORA_NPTS] = {435, 479, 515, 558, 580,
                                        628, 670, 685, 720};       /* :72-73 */

/* params.hpp:75-83: clamp at the ends, linear search, linear blend. */
double ora_interp(double T, const double* xs, const double* ys, int n) {
    if (T <= xs[0]) return ys[0];
    if (T >= xs[n - 1]) return ys[n - 1];
    int i = 1;
    while (T > xs[i]) ++i;
    const double f = (T - xs[i - 1]) / (xs[i] - xs[i - 1]);
    return ys[i - 1] + f * (ys[i] - ys[i - 1]);
}

double ora_k_of(double T)  { return ora_interp(T, T_PTS, K_PTS, ORA_NPTS); }   /* :85 */
double ora_cp_of(double T) { return ora_interp(T, T_PTS, CP_PTS, ORA_NPTS); }  /* :86 */

/* params.hpp:93-100 */
double ora_k_enh(double T, double kmult) {
    if (kmult <= 1.0) return 1.0;
    double s;
    if (T <= T_SOL) s = 0.0;
    else if (T >= T_LIQ) s = 1.0;
    else s = (T - T_SOL) / (T_LIQ - T_SOL);
    return 1.0 + (kmult - 1.0) * s;
}

double ora_k_eff(double T, double kmult) { return ora_k_of(T) * ora_k_enh(T, kmult); }  /* :101-103 */

/* params.hpp:107-111 */
double ora_cp_eff(double T) {
    double c = ora_cp_of(T);
    if (T >= T_SOL && T <= T_LIQ) c += LF / (T_LIQ - T_SOL);
    return c;
}

/* params.hpp:116-121: scan T = 298, 303, ... <= 2200 by repeated += 5. */
double ora_alpha_eff_max(double rho, double kmult) {
    double m = 0.0;
    for (double T = 298.0; T <= 2200.0; T += 5.0) {
        const double a = ora_k_eff(T, kmult) / (rho * ora_cp_of(T));
        if (a > m) m = a;
    }
    return m;
}

/* params.hpp:126-131 */
double ora_dt_cfl_ext(const OraMaterial* m, const OraDomain* d, double kmult) {
    const double a = ora_alpha_eff_max(m->rho, kmult);
    const double dx = ora_dx(d), dy = ora_dy(d);
    return 0.35 / (a * (1.0 / (dx * dx) + 1.0 / (dy * dy)));
}

/* lpbf_step_const.hpp:36-54 */
void ora_make_const(const OraMaterial* m, const OraProcess* p,
                    const OraDomain* d, double dt, StepConst* k) {
    memset(k, 0, sizeof(*k));
    k->nx = d->nx; k->ny = d->ny;
    k->dx = ora_dx(d); k->dy = ora_dy(d);
    k->inv_dx2 = 1.0 / (k->dx * k->dx);
    k->inv_dy2 = 1.0 / (k->dy * k->dy);
    k->inv_r0_sq = 1.0 / (p->r0 * p->r0);
    k->alpha = ora_alpha(m); k->K = m->k; k->rho = m->rho; k->cp = m->cp;
    k->rho_cp = m->rho * m->cp; k->inv_h = 1.0 / p->h; k->h = p->h; k->dt = dt;
    k->peak_S = (2.0 * m->eta * p->P) / (ORA_PI * p->r0 * p->r0);  /* step_task.hpp:68-69 */
    k->eta = m->eta; k->P = p->P;
    k->hConv = p->hConv; k->Tamb = p->Tamb; k->Tsub = p->Tsub; k->dSub = p->dSub;
    k->eps = m->eps; k->radiation = p->radiation;
    k->kmult = 1.0;
    k->q_scale = 1.0;
    k->x_l = 0.0; k->y_l = 0.0;
}

/* ---- the cell update ---------------------------------------------------- */

/* Baseline update of one cell from its centre and four neighbour values
 * (step_task.hpp:93-129, kernels.cu:106-124).  lap_x uses the reference's
 * one-sided edge forms, selected by the caller through xe (-1 left edge,
 * +1 right edge, 0 interior).  src/loss receive dt*S and dt*loss. */
static double cell_base(double T, double Txm, double Txp, double Tym,
                        double Typ, int xe, double xpos, double ypos,
                        const StepConst* k, double* src, double* lossv) {
    double lap_x;
    if (xe < 0)      lap_x = (Txp - T) * k->inv_dx2;
    else if (xe > 0) lap_x = (Txm - T) * k->inv_dx2;
    else             lap_x = (Txp - 2.0 * T + Txm) * k->inv_dx2;
    const double lap_y = (Typ - 2.0 * T + Tym) * k->inv_dy2;
    const double r2 = (xpos - k->x_l) * (xpos - k->x_l)
                    + (ypos - k->y_l) * (ypos - k->y_l);
    const double S = k->q_scale * k->peak_S * exp(-2.0 * r2 * k->inv_r0_sq)
                   / (k->rho_cp * k->h);
    double q_top = k->hConv * (T - k->Tamb);
    if (k->radiation) {
        const double T4 = T * T * T * T;
        const double Ta4 = k->Tamb * k->Tamb * k->Tamb * k->Tamb;
        q_top += k->eps * ORA_SIGMA_SB * (T4 - Ta4);
    }
    const double q_bot = (k->K / k->dSub) * (T - k->Tsub);
    const double loss = (q_top + q_bot) * k->inv_h / k->rho_cp;
    const double dTdt = k->alpha * (lap_x + lap_y) + S - loss;
    if (src) *src = k->dt * S;
    if (lossv) *lossv = k->dt * loss;
    return T + k->dt * dTdt;
}

/* Extended update (step_task.hpp:143-185, kernels.cu:125-147).  Neighbour
 * values are already ghost-substituted by the caller. */
static double cell_ext(double T, double Txm, double Txp, double Tym,
                       double Typ, double xpos, double ypos,
                       const StepConst* k, double* src, double* lossv) {
    const double km = k->kmult;
    const double kCe = ora_k_eff(T, km);
    const double kxm = 0.5 * (kCe + ora_k_eff(Txm, km));
    const double kxp = 0.5 * (kCe + ora_k_eff(Txp, km));
    const double kym = 0.5 * (kCe + ora_k_eff(Tym, km));
    const double kyp = 0.5 * (kCe + ora_k_eff(Typ, km));
    const double div_k_grad =
          (kxp * (Txp - T) - kxm * (T - Txm)) * k->inv_dx2
        + (kyp * (Typ - T) - kym * (T - Tym)) * k->inv_dy2;
    const double r2 = (xpos - k->x_l) * (xpos - k->x_l)
                    + (ypos - k->y_l) * (ypos - k->y_l);
    const double q_in = k->q_scale * k->peak_S * exp(-2.0 * r2 * k->inv_r0_sq);
    double q_top = k->hConv * (T - k->Tamb);
    if (k->radiation) {
        const double T4 = T * T * T * T;
        const double Ta4 = k->Tamb * k->Tamb * k->Tamb * k->Tamb;
        q_top += k->eps * ORA_SIGMA_SB * (T4 - Ta4);
    }
    const double q_bot = (ora_k_of(T) / k->dSub) * (T - k->Tsub);
    const double rhs = div_k_grad + (q_in - q_top - q_bot) * k->inv_h;
    const double rcp = k->rho * ora_cp_eff(T);
    if (src) *src = k->dt * (q_in * k->inv_h) / rcp;
    if (lossv) *lossv = k->dt * ((q_top + q_bot) * k->inv_h) / rcp;
    return T + k->dt * rhs / rcp;
}

/* lpbf_cuda.cu:154-210: full grid, ghost = centre on all edges. */
void ora_step(const double* Tin, double* Tout, const StepConst* k, int phys) {
    ora_step_ex(Tin, Tout, k, phys, LPBF_BC_MIRROR, 0.0, 0.0, NULL);
}

/* ora_step restricted to rows [y0, y1): the same cell arithmetic, so row
 * ranges stepped by separate threads give the ora_step field exactly. */
void ora_step_rows(const double* Tin, double* Tout, const StepConst* k, int phys,
                   int y0, int y1) {
    const int nx = k->nx, ny = k->ny;
    if (y0 < 0) y0 = 0;
    if (y1 > ny) y1 = ny;
    for (int y = y0; y < y1; ++y) {
        const double ypos = (double)y * k->dy;
        for (int x = 0; x < nx; ++x) {
            const long long c = (long long)y * nx + x;
            const double T = Tin[c];
            const double xpos = (double)x * k->dx;
            double src = 0.0, lossv = 0.0;
            const double Txm = (x == 0) ? T : Tin[c - 1];
            const double Txp = (x == nx - 1) ? T : Tin[c + 1];
            const double Tym = (y == 0) ? T : Tin[c - nx];
            const double Typ = (y == ny - 1) ? T : Tin[c + nx];
            if (phys == LPBF_PHYS_BASELINE) {
                const int xe = (x == 0) ? -1 : (x == nx - 1 ? 1 : 0);
                Tout[c] = cell_base(T, Txm, Txp, Tym, Typ, xe, xpos, ypos, k, &src, &lossv);
            } else {
                Tout[c] = cell_ext(T, Txm, Txp, Tym, Typ, xpos, ypos, k, &src, &lossv);
            }
        }
    }
}

/* ora_step_rows in FP32 (baseline physics, mirror edges): the reference
 * expression (step_task.hpp:93-129) evaluated in float storage and
 * arithmetic with expf.  Not a parity function: it is the FP32 CPU baseline
 * the benchmark times beside the FP32 GPU path. */
typedef struct F32Const {
    float inv_dx2, inv_dy2, alpha, dt, dx, dy, x_l, y_l, inv_r0_sq, S0;
    float hConv, Tamb, Tsub, eps_sb, Ta4, Kd, loss_w;
    int rad, nx;
} F32Const;

static F32Const f32_const(const StepConst* k) {
    F32Const c;
    c.inv_dx2 = (float)k->inv_dx2; c.inv_dy2 = (float)k->inv_dy2;
    c.alpha = (float)k->alpha; c.dt = (float)k->dt;
    c.dx = (float)k->dx; c.dy = (float)k->dy;
    c.x_l = (float)k->x_l; c.y_l = (float)k->y_l; c.inv_r0_sq = (float)k->inv_r0_sq;
    c.S0 = (float)(k->q_scale * k->peak_S / (k->rho_cp * k->h));
    c.hConv = (float)k->hConv; c.Tamb = (float)k->Tamb; c.Tsub = (float)k->Tsub;
    c.eps_sb = (float)(k->eps * ORA_SIGMA_SB);
    c.Ta4 = c.Tamb * c.Tamb * c.Tamb * c.Tamb;
    c.Kd = (float)(k->K / k->dSub);
    c.loss_w = (float)(k->inv_h / k->rho_cp);
    c.rad = k->radiation;
    c.nx = k->nx;
    return c;
}

/* One row y of the FP32 step from the row, the row above and the row below
 * (ghost rows already substituted by the caller). */
static void f32_row(const float* row, const float* up, const float* dn, float* out, int y,
                    const F32Const* c) {
    const int nx = c->nx;
    const float ypos = (float)y * c->dy;
    for (int x = 0; x < nx; ++x) {
        const float T = row[x];
        const float xpos = (float)x * c->dx;
        const float Txm = (x == 0) ? T : row[x - 1];
        const float Txp = (x == nx - 1) ? T : row[x + 1];
        float lap_x;
        if (x == 0)           lap_x = (Txp - T) * c->inv_dx2;
        else if (x == nx - 1) lap_x = (Txm - T) * c->inv_dx2;
        else                  lap_x = (Txp - 2.0f * T + Txm) * c->inv_dx2;
        const float lap_y = (dn[x] - 2.0f * T + up[x]) * c->inv_dy2;
        const float r2 = (xpos - c->x_l) * (xpos - c->x_l) + (ypos - c->y_l) * (ypos - c->y_l);
        const float S = c->S0 * expf(-2.0f * r2 * c->inv_r0_sq);
        float q_top = c->hConv * (T - c->Tamb);
        if (c->rad) q_top += c->eps_sb * (T * T * T * T - c->Ta4);
        const float loss = (q_top + c->Kd * (T - c->Tsub)) * c->loss_w;
        out[x] = T + c->dt * (c->alpha * (lap_x + lap_y) + S - loss);
    }
}

void ora_step_rows_f32(const float* Tin, float* Tout, const StepConst* k, int y0, int y1) {
    const int nx = k->nx, ny = k->ny;
    const F32Const c = f32_const(k);
    if (y0 < 0) y0 = 0;
    if (y1 > ny) y1 = ny;
    for (int y = y0; y < y1; ++y) {
        const float* row = Tin + (size_t)y * nx;
        const float* up = (y == 0) ? row : row - nx;
        const float* dn = (y == ny - 1) ? row : row + nx;
        f32_row(row, up, dn, Tout + (size_t)y * nx, y, &c);
    }
}

/* ora_step_tile in FP32 (baseline physics): the tile's halo rows carry the
 * neighbour or mirror rows, so the owned rows equal ora_step_rows_f32's. */
void ora_step_tile_f32(const float* tile, float* out, int owned, int y0, const StepConst* k) {
    const int nx = k->nx, rows = owned + 2;
    const F32Const c = f32_const(k);
    memcpy(out, tile, sizeof(float) * (size_t)nx);
    memcpy(out + (size_t)(rows - 1) * nx, tile + (size_t)(rows - 1) * nx,
           sizeof(float) * (size_t)nx);
    for (int r = 1; r < rows - 1; ++r)
        f32_row(tile + (size_t)r * nx, tile + (size_t)(r - 1) * nx, tile + (size_t)(r + 1) * nx,
                out + (size_t)r * nx, y0 + (r - 1), &c);
}

void ora_step_ex(const double* Tin, double* Tout, const StepConst* k,
                 int phys, int bc, double T_bc, double Tm, StepStats* st) {
    const int nx = k->nx, ny = k->ny;
    StepStats acc;
    acc.tmax = -1.0; acc.arg = (long long)nx * ny; acc.ymin = ny; acc.ymax = -1;
    acc.molten = 0; acc.sum_dT = 0.0; acc.sum_src = 0.0; acc.sum_loss = 0.0;
    for (int y = 0; y < ny; ++y) {
        const double ypos = (double)y * k->dy;
        for (int x = 0; x < nx; ++x) {
            const long long c = (long long)y * nx + x;
            const double T = Tin[c];
            const double xpos = (double)x * k->dx;
            double v, src = 0.0, lossv = 0.0;
            int updated = 1;
            if (bc == LPBF_BC_DIRICHLET &&
                (x == 0 || x == nx - 1 || y == 0 || y == ny - 1)) {
                v = T_bc;
                updated = 0;
            } else {
                const double Txm = (x == 0) ? T : Tin[c - 1];
                const double Txp = (x == nx - 1) ? T : Tin[c + 1];
                const double Tym = (y == 0) ? T : Tin[c - nx];
                const double Typ = (y == ny - 1) ? T : Tin[c + nx];
                if (phys == LPBF_PHYS_BASELINE) {
                    const int xe = (x == 0) ? -1 : (x == nx - 1 ? 1 : 0);
                    v = cell_base(T, Txm, Txp, Tym, Typ, xe, xpos, ypos, k,
                                  &src, &lossv);
                } else {
                    v = cell_ext(T, Txm, Txp, Tym, Typ, xpos, ypos, k,
                                 &src, &lossv);
                }
            }
            Tout[c] = v;
            if (st) {
                if (v > acc.tmax) { acc.tmax = v; acc.arg = c; }
                if (v > Tm) {
                    if (y < acc.ymin) acc.ymin = y;
                    if (y > acc.ymax) acc.ymax = y;
                    acc.molten++;
                }
                if (updated) {
                    acc.sum_dT += v - T;
                    acc.sum_src += src;
                    acc.sum_loss += lossv;
                }
            }
        }
    }
    if (st) *st = acc;
}

/* step_task.hpp:54-206: owned rows 1..owned of a tile, halo rows copied
 * through (:78-84), x edges by the one-sided forms, y neighbours from the
 * tile rows (the halo rows carry the mirror at the domain edge). */
void ora_step_tile(const double* tile, double* out, int owned, int y0,
                   const StepConst* k, int phys) {
    const int nx = k->nx, rows = owned + 2;
    memcpy(out, tile, sizeof(double) * (size_t)nx);
    memcpy(out + (size_t)(rows - 1) * nx, tile + (size_t)(rows - 1) * nx,
           sizeof(double) * (size_t)nx);
    for (int r = 1; r < rows - 1; ++r) {
        const int yg = y0 + (r - 1);
        const double ypos = (double)yg * k->dy;
        const double* row = tile + (size_t)r * nx;
        const double* up = tile + (size_t)(r - 1) * nx;
        const double* dn = tile + (size_t)(r + 1) * nx;
        for (int c = 0; c < nx; ++c) {
            const double T = row[c];
            const double xpos = (double)c * k->dx;
            const double Txm = (c == 0) ? T : row[c - 1];
            const double Txp = (c == nx - 1) ? T : row[c + 1];
            double v;
            if (phys == LPBF_PHYS_BASELINE) {
                const int xe = (c == 0) ? -1 : (c == nx - 1 ? 1 : 0);
                v = cell_base(T, Txm, Txp, up[c], dn[c], xe, xpos, ypos, k,
                              NULL, NULL);
            } else {
                v = cell_ext(T, Txm, Txp, up[c], dn[c], xpos, ypos, k,
                             NULL, NULL);
            }
            out[(size_t)r * nx + c] = v;
        }
    }
}

/* main.cpp:115-123 */
void ora_tile_bounds(int ny, int ntiles, int* y0, int* y1) {
    for (int i = 0; i < ntiles; ++i) {
        y0[i] = (i * ny) / ntiles;
        y1[i] = ((i + 1) * ny) / ntiles - 1;
    }
}

/* main.cpp:125-155: owned rows plus a neighbour row above and below, or a
 * copy of the edge row at the domain top/bottom. */
void ora_make_tile(const double* T, int nx, int ny, int y0, int y1,
                   double* tile) {
    const int owned = y1 - y0 + 1, rows = owned + 2;
    const int top_src = (y0 == 0) ? y0 : (y0 - 1);
    const int bot_src = (y1 == ny - 1) ? y1 : (y1 + 1);
    memcpy(tile, T + (size_t)top_src * nx, sizeof(double) * (size_t)nx);
    for (int r = 0; r < owned; ++r)
        memcpy(tile + (size_t)(r + 1) * nx, T + (size_t)(y0 + r) * nx,
               sizeof(double) * (size_t)nx);
    memcpy(tile + (size_t)(rows - 1) * nx, T + (size_t)bot_src * nx,
           sizeof(double) * (size_t)nx);
}

/* main.cpp:316-348 for one step: build, step, stitch the owned rows. */
void ora_step_tiled(const double* Tin, double* Tout, const StepConst* k,
                    int phys, int ntiles) {
    const int nx = k->nx, ny = k->ny;
    if (ntiles > ny) ntiles = ny;  /* main.cpp:237 */
    int* y0 = (int*)malloc(sizeof(int) * (size_t)ntiles);
    int* y1 = (int*)malloc(sizeof(int) * (size_t)ntiles);
    ora_tile_bounds(ny, ntiles, y0, y1);
    for (int i = 0; i < ntiles; ++i) {
        const int owned = y1[i] - y0[i] + 1;
        const size_t sz = (size_t)(owned + 2) * nx;
        double* tile = (double*)malloc(sizeof(double) * sz);
        double* out = (double*)malloc(sizeof(double) * sz);
        ora_make_tile(Tin, nx, ny, y0[i], y1[i], tile);
        ora_step_tile(tile, out, owned, y0[i], k, phys);
        memcpy(Tout + (size_t)y0[i] * nx, out + nx,
               sizeof(double) * (size_t)owned * nx);
        free(tile); free(out);
    }
    free(y0); free(y1);
}

/* ---- metrics ------------------------------------------------------------ */

/* main.cpp:164-211 */
void ora_measure(const double* T, int nx, int ny, double dy, double Tm,
                 OraMetrics* m) {
    int y_min = ny, y_max = -1, x_hot = 0;
    double Tmax = 0.0;
    for (int y = 0; y < ny; ++y) {
        const double* row = T + (size_t)y * nx;
        int any_molten = 0;
        for (int x = 0; x < nx; ++x) {
            const double v = row[x];
            if (v > Tmax) { Tmax = v; x_hot = x; }
            if (v > Tm) any_molten = 1;
        }
        if (any_molten) {
            if (y < y_min) y_min = y;
            if (y > y_max) y_max = y;
        }
    }
    m->peak_T = Tmax;
    m->x_hot = x_hot;
    m->W_um = (y_max < 0) ? 0.0 : (y_max - y_min + 1) * dy * 1e6;
    m->W_sub_um = 0.0;
    if (Tmax > Tm) {
        int a = -1, b = -1;
        for (int y = 0; y < ny; ++y)
            if (T[(size_t)y * nx + x_hot] > Tm) { if (a < 0) a = y; b = y; }
        if (a >= 0) {
            double y_lo = a * dy, y_hi = b * dy;
            if (a > 0) {
                const double T0 = T[(size_t)(a - 1) * nx + x_hot];
                const double T1 = T[(size_t)a * nx + x_hot];
                y_lo = ((a - 1) + (Tm - T0) / (T1 - T0)) * dy;
            }
            if (b < ny - 1) {
                const double T0 = T[(size_t)b * nx + x_hot];
                const double T1 = T[(size_t)(b + 1) * nx + x_hot];
                y_hi = (b + (Tm - T0) / (T1 - T0)) * dy;
            }
            m->W_sub_um = (y_hi - y_lo) * 1e6;
        }
    }
}

/* gpu_step_task.hpp:380-395 */
double ora_wsub_from_col(const double* col, int ny, double dy, double Tm) {
    int a = -1, b = -1;
    for (int y = 0; y < ny; ++y)
        if (col[y] > Tm) { if (a < 0) a = y; b = y; }
    if (a < 0) return 0.0;
    double y_lo = a * dy, y_hi = b * dy;
    if (a > 0) y_lo = ((a - 1) + (Tm - col[a - 1]) / (col[a] - col[a - 1])) * dy;
    if (b < ny - 1) y_hi = (b + (Tm - col[b]) / (col[b + 1] - col[b])) * dy;
    return (y_hi - y_lo) * 1e6;
}

static void partial_merge(Partial* acc, const Partial* p) {
    if (p->tmax > acc->tmax || (p->tmax == acc->tmax && p->arg < acc->arg)) {
        acc->tmax = p->tmax; acc->arg = p->arg;
    }
    if (p->ymin < acc->ymin) acc->ymin = p->ymin;
    if (p->ymax > acc->ymax) acc->ymax = p->ymax;
}

/* kernels.cu:205-237: block b, thread t visits i = b*256+t, += 256*nblocks;
 * strict '>' keeps the first hit per thread; the block then merges its 256
 * thread records in thread order. */
void ora_reduce_partials(const double* T, int nx, int ny, double Tm,
                         int nblocks, Partial* out) {
    const long long N = (long long)nx * ny;
    for (int b = 0; b < nblocks; ++b) {
        Partial blk = {-1.0, N, ny, -1};
        for (int t = 0; t < 256; ++t) {
            Partial th = {-1.0, N, ny, -1};
            for (long long i = (long long)b * 256 + t; i < N;
                 i += 256LL * nblocks) {
                const double v = T[i];
                if (v > th.tmax) { th.tmax = v; th.arg = i; }
                if (v > Tm) {
                    const int yy = (int)(i / nx);
                    if (yy < th.ymin) th.ymin = yy;
                    if (yy > th.ymax) th.ymax = yy;
                }
            }
            partial_merge(&blk, &th);
        }
        out[b] = blk;
    }
}

/* kernels.cu:248-262 */
void ora_reduce_finalize(const Partial* parts, int nblocks, int nx, int ny,
                         Partial* out) {
    Partial acc = {-1.0, (long long)nx * ny, ny, -1};
    for (int j = 0; j < nblocks; ++j) partial_merge(&acc, &parts[j]);
    *out = acc;
}

/* ---- laser plans -------------------------------------------------------- */

/* hh_gpu_main.cpp:141-157 (= main.cpp:284-286, 351-352, 364-365). */
void ora_build_plan(double V, double dt, double Lx, double Ly, int n_steps,
                    int snap_every, OraTick* plan) {
    double x_laser = 0.20 * Lx, y_laser = 0.50 * Ly;
    const double x_start = x_laser;
    int wrapped = 0;
    for (int s = 0; s < n_steps; ++s) {
        OraTick* tk = &plan[s];
        tk->step = s; tk->t = s * dt;
        tk->x_l = x_laser; tk->y_l = y_laser; tk->q_scale = 1.0;
        x_laser += V * dt;
        if (x_laser > 0.90 * Lx) { x_laser = x_start; wrapped = 1; }
        tk->x_laser_csv = x_laser; tk->wrapped = wrapped;
        tk->dump = (snap_every > 0 && s % snap_every == 0) || (s == n_steps - 1);
    }
}

void ora_track_plan(double V, double dt, double Lx, double Ly, int n_steps,
                    LaserTick* plan) {
    double x_laser = 0.20 * Lx;
    const double y_laser = 0.50 * Ly, x_start = x_laser;
    for (int s = 0; s < n_steps; ++s) {
        plan[s].x_l = x_laser; plan[s].y_l = y_laser; plan[s].q_scale = 1.0;
        x_laser += V * dt;
        if (x_laser > 0.90 * Lx) x_laser = x_start;
    }
}

/* Path: track 0 (+x, beam on), hatch jump (+y, beam off), track 1 (-x), ...
 * The arc length d is a running FP64 sum of V*dt, like the reference law. */
void ora_raster_plan(const OraRaster* r, double dt, int n_steps,
                     LaserTick* plan) {
    const double cycle = r->length + r->hatch;
    const double step_len = r->V * dt;
    double d = 0.0;
    for (int s = 0; s < n_steps; ++s) {
        LaserTick* t = &plan[s];
        int i = (int)floor(d / cycle);
        double u = d - (double)i * cycle;
        if (i >= r->n_tracks || (i == r->n_tracks - 1 && u >= r->length)) {
            const int last = r->n_tracks - 1;
            t->x_l = (last % 2 == 0) ? r->x0 + r->length : r->x0;
            t->y_l = r->y0 + (double)last * r->hatch;
            t->q_scale = 0.0;
        } else if (u < r->length) {
            t->x_l = (i % 2 == 0) ? r->x0 + u : r->x0 + r->length - u;
            t->y_l = r->y0 + (double)i * r->hatch;
            t->q_scale = 1.0;
        } else {
            t->x_l = (i % 2 == 0) ? r->x0 + r->length : r->x0;
            t->y_l = r->y0 + (double)i * r->hatch + (u - r->length);
            t->q_scale = 0.0;
        }
        d += step_len;
    }
}

/* ---- noise -------------------------------------------------------------- */

/* SplitMix64 output function (Steele, Lea & Flood, OOPSLA 2014) applied to
 * the counter idx + (seed+1)*golden; the top 24 bits give u in [0,1). */
double ora_noise(uint64_t seed, uint64_t idx) {
    uint64_t z = idx + (seed + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 40) * (1.0 / 16777216.0) - 0.5;
}

void ora_noise_field(uint64_t seed, double base, double amp, double* T,
                     long long n) {
    for (long long i = 0; i < n; ++i)
        T[i] = base + amp * 2.0 * ora_noise(seed, (uint64_t)i);
}

/* ---- runs --------------------------------------------------------------- */

void ora_run(const OraMaterial* m, const OraProcess* p, const OraDomain* d,
             double dt, int n_steps, int phys, double kmult, int ntiles,
             double* T, double* step_peak, double* step_W, double* step_Wsub,
             double* summary) {
    const int nx = d->nx, ny = d->ny;
    const size_t N = (size_t)nx * ny;
    StepConst k;
    ora_make_const(m, p, d, dt, &k);
    k.kmult = kmult;
    double* A = T;
    double* B = (double*)malloc(sizeof(double) * N);
    double* cur = A;
    double* nxt = B;
    double x_laser = 0.20 * d->Lx;
    const double y_laser = 0.50 * d->Ly, x_start = x_laser;
    double peak_T = p->Tamb, peak_W = 0.0, steady = 0.0;
    int wrapped = 0;
    const double dy = ora_dy(d);
    for (int s = 0; s < n_steps; ++s) {
        k.x_l = x_laser; k.y_l = y_laser;
        if (ntiles > 0) ora_step_tiled(cur, nxt, &k, phys, ntiles);
        else ora_step(cur, nxt, &k, phys);
        double* t = cur; cur = nxt; nxt = t;
        x_laser += p->V * dt;                                   /* main.cpp:351 */
        if (x_laser > 0.90 * d->Lx) { x_laser = x_start; wrapped = 1; }
        OraMetrics mm;
        ora_measure(cur, nx, ny, dy, m->Tm, &mm);
        if (mm.peak_T > peak_T) peak_T = mm.peak_T;
        if (mm.W_um > peak_W) peak_W = mm.W_um;
        const int cruise = !wrapped && x_laser >= 0.45 * d->Lx
                        && x_laser <= 0.85 * d->Lx;
        if (cruise && mm.W_sub_um > steady) steady = mm.W_sub_um;
        if (step_peak) step_peak[s] = mm.peak_T;
        if (step_W) step_W[s] = mm.W_um;
        if (step_Wsub) step_Wsub[s] = mm.W_sub_um;
    }
    if (cur != A) memcpy(A, cur, sizeof(double) * N);
    free(B);
    if (summary) { summary[0] = peak_T; summary[1] = peak_W; summary[2] = steady; }
}

void ora_run_plan(const StepConst* k0, int phys, int bc, double T_bc,
                  double Tm, const LaserTick* plan, int n_steps, double* T,
                  StepStats* stats) {
    const size_t N = (size_t)k0->nx * k0->ny;
    StepConst k = *k0;
    double* B = (double*)malloc(sizeof(double) * N);
    double* cur = T;
    double* nxt = B;
    for (int s = 0; s < n_steps; ++s) {
        k.x_l = plan[s].x_l; k.y_l = plan[s].y_l; k.q_scale = plan[s].q_scale;
        ora_step_ex(cur, nxt, &k, phys, bc, T_bc, Tm, stats ? &stats[s] : NULL);
        double* t = cur; cur = nxt; nxt = t;
    }
    if (cur != T) memcpy(T, cur, sizeof(double) * N);
    free(B);
}
