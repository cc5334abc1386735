// SPDX-License-Identifier: 0BSD
// Per-step laser plans (FP64, host-built, uploaded once to the device).
//
// SingleTrack reproduces the reference laser law (src/main.cpp:284-286 and
// 351-352, precomputed the same way as build_plan in
// src/gpu/hh_gpu_main.cpp:141-157): the source sits at the pre-advance
// position during step s, then x += V*dt and wraps to the start past
// 0.9*Lx.  Raster is new: a bidirectional hatch pattern whose arc length is
// the same running FP64 sum of V*dt, with the beam off (q_scale = 0) on the
// hatch jumps and after the last track.
#pragma once

#include <cmath>
#include <vector>

#include "../lpbf_step_const.h"

namespace lpbf {

// One reference tick: the laser for step s plus the post-advance values the
// reference logs (metrics.csv x_laser_um, the cruise gate, the dump rule).
struct Tick {
    int step;
    double t;
    double x_l, y_l, q_scale;
    double x_laser_csv;
    bool wrapped;
    bool dump;
};

inline std::vector<Tick> single_track_ticks(double V, double dt, double Lx, double Ly,
                                            int n_steps, int snap_every) {
    std::vector<Tick> plan(static_cast<size_t>(n_steps > 0 ? n_steps : 0));
    double x = 0.20 * Lx;
    const double y = 0.50 * Ly, x_start = x;
    bool wrapped = false;
    for (int s = 0; s < n_steps; ++s) {
        Tick& tk = plan[static_cast<size_t>(s)];
        tk.step = s;
        tk.t = s * dt;
        tk.x_l = x;
        tk.y_l = y;
        tk.q_scale = 1.0;
        x += V * dt;
        if (x > 0.90 * Lx) {
            x = x_start;
            wrapped = true;
        }
        tk.x_laser_csv = x;
        tk.wrapped = wrapped;
        tk.dump = (snap_every > 0 && s % snap_every == 0) || (s == n_steps - 1);
    }
    return plan;
}

inline std::vector<LaserTick> to_laser(std::vector<Tick> const& ticks) {
    std::vector<LaserTick> out(ticks.size());
    for (size_t i = 0; i < ticks.size(); ++i) out[i] = {ticks[i].x_l, ticks[i].y_l, ticks[i].q_scale};
    return out;
}

struct Raster {
    double x0, y0;   // start of track 0 [m]
    double length;   // track length [m]
    double hatch;    // track spacing [m]
    int n_tracks;
    double V;        // scan and jump speed [m/s]
};

inline std::vector<LaserTick> raster_plan(Raster const& r, double dt, int n_steps) {
    std::vector<LaserTick> plan(static_cast<size_t>(n_steps > 0 ? n_steps : 0));
    const double cycle = r.length + r.hatch;
    const double step_len = r.V * dt;
    const int last = r.n_tracks - 1;
    double d = 0.0;
    for (int s = 0; s < n_steps; ++s) {
        LaserTick& t = plan[static_cast<size_t>(s)];
        const int i = static_cast<int>(std::floor(d / cycle));
        const double u = d - static_cast<double>(i) * cycle;
        if (i >= r.n_tracks || (i == last && u >= r.length)) {
            t.x_l = (last % 2 == 0) ? r.x0 + r.length : r.x0;
            t.y_l = r.y0 + static_cast<double>(last) * r.hatch;
            t.q_scale = 0.0;
        } else if (u < r.length) {
            t.x_l = (i % 2 == 0) ? r.x0 + u : r.x0 + r.length - u;
            t.y_l = r.y0 + static_cast<double>(i) * r.hatch;
            t.q_scale = 1.0;
        } else {
            t.x_l = (i % 2 == 0) ? r.x0 + r.length : r.x0;
            t.y_l = r.y0 + static_cast<double>(i) * r.hatch + (u - r.length);
            t.q_scale = 0.0;
        }
        d += step_len;
    }
    return plan;
}

// A raster of n_tracks spending n_steps of travel at speed V, centred in an
// Lx x Ly plate: track length = (V*dt*n_steps - (n_tracks-1)*hatch)/n_tracks.
inline Raster centred_raster(double Lx, double Ly, double V, double dt, int n_steps,
                             int n_tracks, double hatch) {
    Raster r;
    r.n_tracks = n_tracks;
    r.hatch = hatch;
    r.V = V;
    r.length = (V * dt * n_steps - (n_tracks - 1) * hatch) / n_tracks;
    r.x0 = 0.5 * (Lx - r.length);
    r.y0 = 0.5 * Ly - 0.5 * (n_tracks - 1) * hatch;
    return r;
}

}  // namespace lpbf
