// SPDX-License-Identifier: 0BSD
// One row strip of the global field with one halo row above and below:
// the unit of work of the step-task interface.  Same members and accessors
// as the reference include/tile.hpp:10-33.
#pragma once

#include <vector>

namespace lpbf {

struct HaloTile {
    int idx = 0;
    int y0 = 0;   // first owned global row
    int y1 = 0;   // last owned global row (inclusive)
    int nx = 0;   // columns

    int step = 0;
    double t = 0.0;
    double x_laser = 0.0;
    double y_laser = 0.0;
    double q_scale = 1.0;   // 0 while the beam is off

    std::vector<double> data;   // (owned_rows()+2) * nx, row-major, halos at both ends

    int owned_rows() const noexcept { return y1 - y0 + 1; }
    int total_rows() const noexcept { return owned_rows() + 2; }
    double& at(int r, int c) noexcept { return data[static_cast<size_t>(r) * nx + c]; }
    double at(int r, int c) const noexcept { return data[static_cast<size_t>(r) * nx + c]; }
};

}  // namespace lpbf
