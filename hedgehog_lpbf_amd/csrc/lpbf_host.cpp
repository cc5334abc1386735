// SPDX-License-Identifier: 0BSD
// Host-side C ABI: configuration arithmetic and laser plans.
#include "lpbf_host.h"

#include <cstring>
#include <vector>

#include "lpbf/params.hpp"
#include "lpbf/plan.hpp"

namespace {

void fill(const double* mat, const double* proc, int radiation, double Lx, double Ly, int nx,
          int ny, lpbf::Material& m, lpbf::Process& p, lpbf::Domain& d) {
    if (mat) {
        m.k = mat[0]; m.rho = mat[1]; m.cp = mat[2]; m.Tm = mat[3]; m.eps = mat[4]; m.eta = mat[5];
    }
    if (proc) {
        p.P = proc[0]; p.V = proc[1]; p.r0 = proc[2]; p.h = proc[3]; p.hConv = proc[4];
        p.dSub = proc[5]; p.Tamb = proc[6]; p.Tsub = proc[7];
    }
    p.radiation = radiation != 0;
    d.Lx = Lx; d.Ly = Ly; d.nx = nx; d.ny = ny;
}

}  // namespace

extern "C" {

void lpbf_make_const(const double* mat, const double* proc, int radiation, double Lx, double Ly,
                     int nx, int ny, double dt, StepConst* out) {
    lpbf::Material m; lpbf::Process p; lpbf::Domain d;
    fill(mat, proc, radiation, Lx, Ly, nx, ny, m, p, d);
    *out = make_const(m, p, d, dt);
}

double lpbf_dt_cfl(const double* mat, double Lx, double Ly, int nx, int ny) {
    lpbf::Material m; lpbf::Process p; lpbf::Domain d;
    fill(mat, nullptr, 1, Lx, Ly, nx, ny, m, p, d);
    return lpbf::dt_cfl(m, d);
}

double lpbf_dt_cfl_ext(const double* mat, double Lx, double Ly, int nx, int ny, double kmult) {
    lpbf::Material m; lpbf::Process p; lpbf::Domain d;
    fill(mat, nullptr, 1, Lx, Ly, nx, ny, m, p, d);
    return lpbf::dt_cfl_ext(m, d, kmult);
}

void lpbf_plan_single_track(double V, double dt, double Lx, double Ly, int n_steps,
                            LaserTick* out, double* x_csv, int* wrapped) {
    const std::vector<lpbf::Tick> t = lpbf::single_track_ticks(V, dt, Lx, Ly, n_steps, 0);
    for (size_t i = 0; i < t.size(); ++i) {
        out[i] = {t[i].x_l, t[i].y_l, t[i].q_scale};
        if (x_csv) x_csv[i] = t[i].x_laser_csv;
        if (wrapped) wrapped[i] = t[i].wrapped ? 1 : 0;
    }
}

void lpbf_plan_raster(double x0, double y0, double length, double hatch, int n_tracks, double V,
                      double dt, int n_steps, LaserTick* out) {
    const lpbf::Raster r{x0, y0, length, hatch, n_tracks, V};
    const std::vector<LaserTick> p = lpbf::raster_plan(r, dt, n_steps);
    if (!p.empty()) std::memcpy(out, p.data(), p.size() * sizeof(LaserTick));
}

}  // extern "C"
