// SPDX-License-Identifier: 0BSD
//
// TEST INFRASTRUCTURE ONLY.  Exposes the reference's own header-only
// arithmetic (include/params.hpp and include/lpbf_step_const.hpp of
// hedgehog_lpbf, compiled where they lie) through a C ABI, so the tests can
// check the oracle's constants, dt rules and IN718 tables bit for bit
// against the reference.  Built by oracle/Makefile into oracle/_ref/ only
// when the reference tree is present; nothing here is part of the product.
#include <cstring>

#include "params.hpp"
#include "lpbf_step_const.hpp"

namespace {

void fill(const double* mat, const double* proc, int radiation, double Lx,
          double Ly, int nx, int ny, lpbf::Material& m, lpbf::Process& p,
          lpbf::Domain& d) {
    m.k = mat[0]; m.rho = mat[1]; m.cp = mat[2]; m.Tm = mat[3];
    m.eps = mat[4]; m.eta = mat[5];
    p.P = proc[0]; p.V = proc[1]; p.r0 = proc[2]; p.h = proc[3];
    p.hConv = proc[4]; p.dSub = proc[5]; p.Tamb = proc[6]; p.Tsub = proc[7];
    p.radiation = radiation != 0;
    d.Lx = Lx; d.Ly = Ly; d.nx = nx; d.ny = ny;
}

}  // namespace

extern "C" {

int ref_sizeof_stepconst() { return (int)sizeof(StepConst); }
int ref_sizeof_partial() { return (int)sizeof(Partial); }

// mat = {k, rho, cp, Tm, eps, eta}; proc = {P, V, r0, h, hConv, dSub, Tamb, Tsub}
void ref_make_const(const double* mat, const double* proc, int radiation,
                    double Lx, double Ly, int nx, int ny, double dt,
                    void* out) {
    lpbf::Material m; lpbf::Process p; lpbf::Domain d;
    fill(mat, proc, radiation, Lx, Ly, nx, ny, m, p, d);
    StepConst k = make_const(m, p, d, dt);
    std::memcpy(out, &k, sizeof(k));
}

double ref_dt_cfl(const double* mat, double Lx, double Ly, int nx, int ny) {
    lpbf::Material m; lpbf::Domain d;
    m.k = mat[0]; m.rho = mat[1]; m.cp = mat[2];
    d.Lx = Lx; d.Ly = Ly; d.nx = nx; d.ny = ny;
    return lpbf::dt_cfl(m, d);
}

double ref_dt_cfl_ext(const double* mat, double Lx, double Ly, int nx, int ny,
                      double kmult) {
    lpbf::Material m; lpbf::Domain d;
    m.k = mat[0]; m.rho = mat[1]; m.cp = mat[2];
    d.Lx = Lx; d.Ly = Ly; d.nx = nx; d.ny = ny;
    return lpbf::dt_cfl_ext(m, d, kmult);
}

double ref_k_of(double T) { return lpbf::in718::k_of(T); }
double ref_cp_of(double T) { return lpbf::in718::cp_of(T); }
double ref_k_enh(double T, double kmult) { return lpbf::in718::k_enh(T, kmult); }
double ref_k_eff(double T, double kmult) { return lpbf::in718::k_eff(T, kmult); }
double ref_cp_eff(double T) { return lpbf::in718::cp_eff(T); }
double ref_alpha_eff_max(double rho, double kmult) {
    return lpbf::in718::alpha_eff_max(rho, kmult);
}

}  // extern "C"
