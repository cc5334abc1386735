// SPDX-License-Identifier: 0BSD
// lpbf_tiles: the reference CPU solver's host loop (src/main.cpp:217-467) on
// the GPU-backed step task.  It exercises the step-task drop-in end to end:
// row-strip tiles with mirrored halos are built on the host each step,
// pushed through lpbf::Graph to StepTileTask workers (each stepping its tile
// on the GPU), stitched back, the laser advanced by the reference law and
// the global metrics measured; metrics.csv / meta.json follow the reference
// formats (main.cpp:281-282, 364-379, 428-467), and perstep.csv carries the
// reference's per-step phase timers (main.cpp:298-310, 382-401).
//
//   lpbf_tiles --case=0 --threads=4 --tiles=8 --tend=1.5e-3 --snap-every=200
//              [--physics=base|ext] [--kmult=F] [--P= --V= --r0= --eta=
//               --dsub= --hconv=] [--out=DIR]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lpbf/params.hpp"
#include "lpbf/plan.hpp"
#include "lpbf/step_task.hpp"
#include "lpbf/tile.hpp"

using namespace lpbf;

namespace {

struct Options {
    std::string case_id = "0";
    int threads = 4;
    int tiles = 8;
    double t_end = 1.5e-3;
    int snap_every = 100;
    std::string physics = "base";
    std::string out;
    double kmult = 1.0;
    std::map<std::string, double> overrides;
};

Options parse(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        const auto eq = a.find('=');
        const std::string key = a.substr(0, eq);
        const std::string val = eq == std::string::npos ? "" : a.substr(eq + 1);
        if (key == "--case") o.case_id = val;
        else if (key == "--threads") o.threads = std::atoi(val.c_str());
        else if (key == "--tiles") o.tiles = std::atoi(val.c_str());
        else if (key == "--tend") o.t_end = std::atof(val.c_str());
        else if (key == "--snap-every") o.snap_every = std::atoi(val.c_str());
        else if (key == "--physics") o.physics = val;
        else if (key == "--out") o.out = val;
        else if (key == "--kmult") o.kmult = std::atof(val.c_str());
        else if (key == "--P" || key == "--V" || key == "--r0" || key == "--eta" ||
                 key == "--dsub" || key == "--hconv")
            o.overrides[key.substr(2)] = std::atof(val.c_str());
        else {
            std::cerr << "unknown option " << a << "\n";
            std::exit(1);
        }
    }
    return o;
}

bool set_case(std::string const& id, Process& pr) {
    struct C { const char* id; double P, V, r0; };
    static const C table[] = {{"0", 285.0, 0.96, 33.5e-6},   {"1.1", 285.0, 0.96, 24.5e-6},
                              {"1.2", 285.0, 0.96, 41.0e-6}, {"2.1", 285.0, 1.20, 33.5e-6},
                              {"2.2", 285.0, 0.80, 33.5e-6}, {"3.1", 325.0, 0.96, 33.5e-6},
                              {"3.2", 245.0, 0.96, 33.5e-6}};
    for (auto const& c : table)
        if (id == c.id) {
            pr.P = c.P; pr.V = c.V; pr.r0 = c.r0;
            return true;
        }
    return false;
}

struct Metrics { double peak_T, W_um, W_sub_um; };

// Global peak T (first in row-major order), molten row extent and the
// sub-grid width through the hottest column (main.cpp:164-211).
Metrics measure(std::vector<double> const& T, Domain const& d, double Tm) {
    int ylo = d.ny, yhi = -1, xh = 0;
    double tmax = 0.0;
    for (int y = 0; y < d.ny; ++y) {
        bool molten = false;
        const double* row = &T[static_cast<size_t>(y) * d.nx];
        for (int x = 0; x < d.nx; ++x) {
            if (row[x] > tmax) { tmax = row[x]; xh = x; }
            molten |= row[x] > Tm;
        }
        if (molten) { ylo = std::min(ylo, y); yhi = y; }
    }
    Metrics m{tmax, yhi < 0 ? 0.0 : (yhi - ylo + 1) * d.dy() * 1e6, 0.0};
    if (tmax > Tm) {
        auto col = [&](int y) { return T[static_cast<size_t>(y) * d.nx + xh]; };
        int a = -1, b = -1;
        for (int y = 0; y < d.ny; ++y)
            if (col(y) > Tm) { if (a < 0) a = y; b = y; }
        const double dy = d.dy();
        double lo = a * dy, hi = b * dy;
        if (a > 0) lo = ((a - 1) + (Tm - col(a - 1)) / (col(a) - col(a - 1))) * dy;
        if (b < d.ny - 1) hi = (b + (Tm - col(b)) / (col(b + 1) - col(b))) * dy;
        m.W_sub_um = (hi - lo) * 1e6;
    }
    return m;
}

}  // namespace

int main(int argc, char** argv) {
    const Options o = parse(argc, argv);
    Material mat;
    Process pr;
    Domain dom;
    if (!set_case(o.case_id, pr)) {
        std::cerr << "unknown case '" << o.case_id << "'\n";
        return 1;
    }
    auto ov = [&](const char* k, double& dst, bool allow_zero = false) {
        auto it = o.overrides.find(k);
        if (it != o.overrides.end() && (it->second > 0 || (allow_zero && it->second >= 0)))
            dst = it->second;
    };
    ov("P", pr.P); ov("V", pr.V); ov("r0", pr.r0); ov("eta", mat.eta); ov("dsub", pr.dSub);
    ov("hconv", pr.hConv, true);

    const Physics phys = o.physics == "ext" ? Physics::Extended : Physics::Baseline;
    const double dt = phys == Physics::Extended ? dt_cfl_ext(mat, dom, o.kmult) : dt_cfl(mat, dom);
    const int n_steps = static_cast<int>(o.t_end / dt);
    const int ntiles = std::min(o.tiles, dom.ny);
    const std::string dir = o.out.empty()
        ? "snapshots_case" + o.case_id + (phys == Physics::Extended ? "_ext" : "") : o.out;
    std::filesystem::create_directories(dir);

    auto prof = std::make_shared<StepProfile>();
    Graph<HaloTile, HaloTile> graph("LPBF-2D");
    auto task = std::make_shared<StepTileTask>("Step", o.threads, mat, pr, dom, dt, prof, phys,
                                               o.kmult);
    graph.inputs(task);
    graph.outputs(task);
    graph.executeGraph();

    std::vector<double> T(static_cast<size_t>(dom.nx) * dom.ny, pr.Tamb);
    std::vector<int> y0(ntiles), y1(ntiles);
    for (int i = 0; i < ntiles; ++i) {
        y0[i] = i * dom.ny / ntiles;
        y1[i] = (i + 1) * dom.ny / ntiles - 1;
    }
    const std::vector<Tick> plan =
        single_track_ticks(pr.V, dt, dom.Lx, dom.Ly, n_steps, o.snap_every);

    std::ofstream csv(dir + "/metrics.csv");
    csv << "step,t_us,x_laser_um,peakT_K,W_um\n";
    // per-step phase timers, the columns of the reference driver's
    // perstep.csv (src/main.cpp:298-310, 382-401): build the tiles, push
    // them, wait for results, stitch them, advance the laser, measure, dump
    std::ofstream perstep(dir + "/perstep.csv");
    perstep << "step,build_us,push_us,wait_us,stitch_us,advance_us,metrics_us,snapshot_us,"
               "total_us\n";
    using clk = std::chrono::steady_clock;
    auto us = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::micro>(b - a).count();
    };
    enum { kBuild, kPush, kWait, kStitch, kAdvance, kMetrics, kSnapshot, kPhases };
    double phase_us[kPhases] = {};
    double peak_T = pr.Tamb, peak_W = 0.0, steady = 0.0;
    const auto wall0 = clk::now();
    for (int s = 0; s < n_steps; ++s) {
        const Tick& tk = plan[static_cast<size_t>(s)];
        const auto tb0 = clk::now();
        std::vector<std::shared_ptr<HaloTile>> tiles(static_cast<size_t>(ntiles));
        for (int i = 0; i < ntiles; ++i) {
            auto t = std::make_shared<HaloTile>();
            t->idx = i; t->y0 = y0[i]; t->y1 = y1[i]; t->nx = dom.nx;
            t->step = s; t->t = tk.t; t->x_laser = tk.x_l; t->y_laser = tk.y_l;
            t->q_scale = tk.q_scale;
            t->data.resize(static_cast<size_t>(t->total_rows()) * dom.nx);
            const int top = y0[i] == 0 ? 0 : y0[i] - 1;
            const int bot = y1[i] == dom.ny - 1 ? y1[i] : y1[i] + 1;
            auto row = [&](int y) { return T.begin() + static_cast<long>(y) * dom.nx; };
            std::copy(row(top), row(top) + dom.nx, t->data.begin());
            std::copy(row(y0[i]), row(y1[i] + 1), t->data.begin() + dom.nx);
            std::copy(row(bot), row(bot) + dom.nx,
                      t->data.begin() + static_cast<long>(t->total_rows() - 1) * dom.nx);
            tiles[static_cast<size_t>(i)] = std::move(t);
        }
        const auto tb1 = clk::now();
        for (auto& t : tiles) graph.pushData(t);
        tiles.clear();
        const auto tp1 = clk::now();
        double wait_us = 0.0, stitch_us = 0.0;
        for (int got = 0; got < ntiles; ++got) {
            const auto tw0 = clk::now();
            auto r = graph.getBlockingResult();
            const auto tw1 = clk::now();
            wait_us += us(tw0, tw1);
            if (!r) break;
            auto t = std::get<std::shared_ptr<HaloTile>>(*r);
            std::copy(t->data.begin() + dom.nx,
                      t->data.begin() + static_cast<long>(t->owned_rows() + 1) * dom.nx,
                      T.begin() + static_cast<long>(t->y0) * dom.nx);
            stitch_us += us(tw1, clk::now());
        }
        // the laser law is precomputed per step (plan.hpp); advancing is
        // taking the post-advance position and the cruise gate of this step
        const auto ta0 = clk::now();
        const bool cruise = !tk.wrapped && tk.x_laser_csv >= 0.45 * dom.Lx &&
                            tk.x_laser_csv <= 0.85 * dom.Lx;
        const auto ta1 = clk::now();
        const Metrics m = measure(T, dom, mat.Tm);
        peak_T = std::max(peak_T, m.peak_T);
        peak_W = std::max(peak_W, m.W_um);
        if (cruise && m.W_sub_um > steady) steady = m.W_sub_um;
        const auto tm1 = clk::now();
        if (tk.dump) {
            std::ofstream bin(dir + "/" + [&] {
                char b[32];
                std::snprintf(b, sizeof b, "T_step_%06d.bin", s);
                return std::string(b);
            }(), std::ios::binary);
            bin.write(reinterpret_cast<const char*>(T.data()),
                      static_cast<std::streamsize>(T.size() * sizeof(double)));
            csv << s << ',' << tk.t * 1e6 << ',' << tk.x_laser_csv * 1e6 << ',' << m.peak_T << ','
                << m.W_um << '\n';
        }
        const auto td1 = clk::now();
        const double step_us[kPhases] = {us(tb0, tb1), us(tb1, tp1), wait_us, stitch_us,
                                         us(ta0, ta1), us(ta1, tm1), us(tm1, td1)};
        perstep << s;
        for (int q = 0; q < kPhases; ++q) {
            perstep << ',' << step_us[q];
            phase_us[q] += step_us[q];
        }
        perstep << ',' << us(tb0, td1) << '\n';
    }
    graph.finishPushingData();
    graph.waitForTermination();
    const auto wall_ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                             clk::now() - wall0).count();

    // meta.json with the reference driver's keys (src/main.cpp:428-467)
    std::ofstream meta(dir + "/meta.json");
    meta << "{\n"
         << "  \"case\":     \"" << o.case_id << "\",\n"
         << "  \"physics\":  \"" << o.physics << "\",\n"
         << "  \"kmult\":    " << o.kmult << ",\n"
         << "  \"eta\":      " << mat.eta << ",\n"
         << "  \"dsub_m\":   " << pr.dSub << ",\n"
         << "  \"hconv\":    " << pr.hConv << ",\n"
         << "  \"P_W\":      " << pr.P << ",\n"
         << "  \"V_mps\":    " << pr.V << ",\n"
         << "  \"r0_m\":     " << pr.r0 << ",\n"
         << "  \"nx\":       " << dom.nx << ",\n"
         << "  \"ny\":       " << dom.ny << ",\n"
         << "  \"Lx_m\":     " << dom.Lx << ",\n"
         << "  \"Ly_m\":     " << dom.Ly << ",\n"
         << "  \"Tm_K\":     " << mat.Tm << ",\n"
         << "  \"dt_s\":     " << dt << ",\n"
         << "  \"n_steps\":  " << n_steps << ",\n"
         << "  \"ntiles\":   " << ntiles << ",\n"
         << "  \"threads\":  " << o.threads << ",\n"
         << "  \"peak_T_K\": " << peak_T << ",\n"
         << "  \"peak_W_um\":" << peak_W << ",\n"
         << "  \"steady_W_sub_um\": " << steady << ",\n"
         << "  \"wall_ms\":  " << wall_ms << ",\n"
         << "  \"phases_ms\": {\n";
    static const char* const names[kPhases] = {"build", "push", "wait", "stitch", "advance",
                                               "metrics", "snapshot"};
    for (int q = 0; q < kPhases; ++q)
        meta << "    \"" << names[q] << "\": " << phase_us[q] / 1e3
             << (q + 1 < kPhases ? ",\n" : "\n");
    meta << "  },\n"
         << "  \"task_ms\": {\n"
         << "    \"compute_total_cpu\":  " << prof->compute_ns.load() / 1e6 << ",\n"
         << "    \"alloc_total_cpu\":    " << prof->alloc_ns.load() / 1e6 << ",\n"
         << "    \"stencil_total_cpu\":  " << prof->stencil_ns.load() / 1e6 << ",\n"
         << "    \"n_calls\":             " << prof->n_calls.load() << "\n"
         << "  }\n"
         << "}\n";
    std::cout << "peak T " << peak_T << " K, peak W " << peak_W << " um, steady W_sub " << steady
              << " um, " << n_steps << " steps in " << wall_ms << " ms -> " << dir << "\n";
    return 0;
}
