// SPDX-License-Identifier: 0BSD
// lpbf_gpu: the device-resident single-GPU solver, with the command line and
// outputs of the reference Hedgehog+CUDA driver (src/gpu/hh_gpu_main.cpp):
//
//   single run   metrics.csv (dump rows), T_step_%06d.bin snapshots written by
//                a pool of --threads writer threads, meta.json (:375-407);
//   --bench      wall time and steps/s, no files (:409-416);
//   --quiet      one RESULT line (:201-210, 321-331);
//   --manifest   a CSV of runs, one RESULT line each, in one process (:234-290);
//                --jobs=N runs N rows at once (one stream each), --ensemble
//                steps the rows of each physics together (lpbf::Ensemble: in
//                FP64 one launch per step for all of them, in FP32 the
//                batched kernel, up to --depth steps per launch).
//
// The field stays on the GPU for the whole run (lpbf::DeviceRun); the host
// synchronises once per --batch steps.  --precision=32 selects the FP32
// temporally blocked path (--depth steps per launch, --rows band height);
// --bc=dirichlet holds the edges at --Tbc.  --mode/--fused/--pinned are
// accepted for command-line compatibility; the batched, fused, pinned path
// is the only one.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lpbf/device_run.hpp"
#include "lpbf/ensemble.hpp"
#include "lpbf/params.hpp"
#include "lpbf/plan.hpp"
#include "lpbf/task_graph.hpp"

using namespace lpbf;

namespace {

struct CLI {
    std::string case_id = "0";
    int n_threads = 4;
    double t_end = 1.5e-3;
    int snap_every = 100;
    std::string physics = "base";
    std::string out_dir;
    double P_ovr = -1, V_ovr = -1, r0_ovr = -1;
    double eta_ovr = -1, dsub_ovr = -1, hconv_ovr = -1;
    double kmult = 1.0;
    int device = 0;
    int batch = 64;
    int mode = 2;
    bool fused = false, pinned = false;
    bool do_bench = false, quiet = false;
    std::string manifest;
    int precision = 64, depth = 4, rows = 0;
    int jobs = 1;                  // manifest rows run concurrently (one stream each)
    bool ensemble = false;         // manifest rows stepped together, one launch per step
    int f64_depth = 1;             // FP64 steps per launch (blocked kernel above 1)
    bool cruise_cols = false;      // FP32: hot column at every cruise step
    std::string bc = "mirror";
    double T_bc = 300.0;
};

[[noreturn]] void usage(int rc) {
    std::cout
        << "Usage: lpbf_gpu [options]   (device-resident 2D z-averaged solver, MI355X)\n"
           "  --case=ID            0,1.1,1.2,2.1,2.2,3.1,3.2 (default 0)\n"
           "  --physics=base|ext   constant props | Mills k(T),cp(T)+latent\n"
           "  --P=W --V=M/S --r0=M     process overrides (SI)\n"
           "  --eta=F --dsub=M --hconv=W/M2K   model overrides\n"
           "  --kmult=F            melt-pool conductivity enhancement\n"
           "  --out=DIR            output directory override\n"
           "  --tend=SECS          physical end-time (default 1.5e-3)\n"
           "  --snap-every=N       snapshot stride (0: final only)\n"
           "  --threads=N          snapshot writer threads\n"
           "  --batch=K            steps per host synchronisation (default 64)\n"
           "  --precision=64|32    FP64 parity kernels | FP32 temporally blocked path\n"
           "  --depth=K --rows=R   FP32 steps per launch (1..4) and band height\n"
           "  --f64-depth=K        FP64 steps per launch (1..4); above 1 the temporally\n"
           "                       blocked kernel between cruise/dump steps (large plates)\n"
           "  --cruise-cols        FP32: end launches at cruise steps too, so W_sub is\n"
           "                       sampled at every cruise step as in the FP64 path\n"
           "  --bc=mirror|dirichlet --Tbc=K   boundary condition\n"
           "  --device=N           HIP device (default 0)\n"
           "  --bench              print wall ms + steps/s (no file output)\n"
           "  --quiet              emit one compact RESULT csv line (no disk output)\n"
           "  --manifest=FILE      loop a runs CSV in-process\n"
           "  --jobs=N             run N manifest rows at once, each on its own stream\n"
           "                       (RESULT lines still in manifest order)\n"
           "  --ensemble           step all manifest rows of one physics together: FP64\n"
           "                       one launch per step, FP32 (--precision=32, baseline\n"
           "                       physics) up to --depth steps per launch; wall_ms is\n"
           "                       the group's\n";
    std::exit(rc);
}

CLI parse_cli(int argc, char** argv) {
    CLI c;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        const auto eq = a.find('=');
        const std::string key = a.substr(0, eq);
        const std::string v = eq == std::string::npos ? std::string() : a.substr(eq + 1);
        const double f = std::atof(v.c_str());
        const int n = std::atoi(v.c_str());
        if (key == "--case") c.case_id = v;
        else if (key == "--threads") c.n_threads = n;
        else if (key == "--tend") c.t_end = f;
        else if (key == "--snap-every") c.snap_every = n;
        else if (key == "--physics") c.physics = v;
        else if (key == "--out") c.out_dir = v;
        else if (key == "--P") c.P_ovr = f;
        else if (key == "--V") c.V_ovr = f;
        else if (key == "--r0") c.r0_ovr = f;
        else if (key == "--eta") c.eta_ovr = f;
        else if (key == "--dsub") c.dsub_ovr = f;
        else if (key == "--hconv") c.hconv_ovr = f;
        else if (key == "--kmult") c.kmult = f;
        else if (key == "--device") c.device = n;
        else if (key == "--batch") c.batch = n;
        else if (key == "--mode") c.mode = n;
        else if (a == "--fused") c.fused = true;
        else if (a == "--pinned") c.pinned = true;
        else if (a == "--bench") c.do_bench = true;
        else if (a == "--quiet") c.quiet = true;
        else if (key == "--manifest") c.manifest = v;
        else if (key == "--precision") c.precision = n;
        else if (key == "--depth") c.depth = n;
        else if (key == "--rows") c.rows = n;
        else if (key == "--bc") c.bc = v;
        else if (key == "--Tbc") c.T_bc = f;
        else if (key == "--jobs") c.jobs = n < 1 ? 1 : n;
        else if (a == "--ensemble") c.ensemble = true;
        else if (key == "--f64-depth") c.f64_depth = n;
        else if (a == "--cruise-cols") c.cruise_cols = true;
        else if (key == "--check" || key == "--check-rtol" || key == "--check-atol") {
            // the reference binary runs its CPU stepper here (lpbf_cuda.cu:485-514);
            // this binary carries no CPU solver
            std::cerr << a << ": the CPU-vs-GPU report is `python -m tests.check_run` "
                                "(same run options, --check-rtol/--check-atol)\n";
            std::exit(1);
        }
        else if (a == "--help" || a == "-h") usage(0);
        else {
            std::cerr << "unknown option " << a << "\n";
            usage(1);
        }
    }
    if (c.precision != 64 && c.precision != 32) {
        std::cerr << "--precision must be 64 or 32\n";
        std::exit(1);
    }
    if (c.bc != "mirror" && c.bc != "dirichlet") {
        std::cerr << "--bc must be mirror or dirichlet\n";
        std::exit(1);
    }
    return c;
}

bool apply_case(std::string const& id, Process& pr) {
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

RunOptions run_options(CLI const& c) {
    RunOptions o;
    o.precision = c.precision;
    o.bc = c.bc == "dirichlet" ? LPBF_BC_DIRICHLET : LPBF_BC_MIRROR;
    o.T_bc = c.T_bc;
    o.batch = c.batch < 1 ? 1 : c.batch;
    o.depth = c.depth;
    o.rows = c.rows;
    o.f64_depth = c.f64_depth;
    o.cruise_cols = c.cruise_cols;
    return o;
}

struct Snapshot {
    int step;
    std::vector<double> field;
};

// Writes T_step_%06d.bin off the stepping thread (reference
// SnapshotWriterTask, gpu_step_task.hpp:417-440).
class SnapshotWriter : public AbstractTask<Snapshot, int> {
public:
    SnapshotWriter(std::size_t n, std::string dir)
        : AbstractTask<Snapshot, int>("SnapshotWriter", n), dir_(std::move(dir)) {}
    void execute(std::shared_ptr<Snapshot> s) override {
        char name[64];
        std::snprintf(name, sizeof name, "/T_step_%06d.bin", s->step);
        std::ofstream out(dir_ + name, std::ios::binary);
        out.write(reinterpret_cast<const char*>(s->field.data()),
                  static_cast<std::streamsize>(s->field.size() * sizeof(double)));
        if (!out) throw std::runtime_error("cannot write snapshot " + dir_ + name);
    }
    std::shared_ptr<AbstractTask<Snapshot, int>> copy() override {
        return std::make_shared<SnapshotWriter>(numberThreads(), dir_);
    }

private:
    std::string dir_;
};

struct RunSpec {
    std::string run_id, cond_id, draw_id, config, physics;
    double P = 0, V = 0, r0 = 0, eta = 0, dsub_m = 0, kmult = 1.0;
};
struct RowResult { double steady_W = 0, peak_T = 0, wall_ms = 0; long n_steps = 0; };

// The run a manifest row describes (reference run_one_row,
// hh_gpu_main.cpp:170-185): default plate and material, the row's process
// values, the CFL step of its physics, a single track to t_end.
struct RowSetup {
    Domain dom;
    StepConst k;
    int phys;
    double Tm, Tamb;
    std::vector<Tick> plan;
};

RowSetup row_setup(RunSpec const& rs, double t_end) {
    Material mat;
    Process pr;
    Domain dom;
    pr.P = rs.P; pr.V = rs.V; pr.r0 = rs.r0;
    mat.eta = rs.eta; pr.dSub = rs.dsub_m;
    const Physics phys = rs.physics == "ext" ? Physics::Extended : Physics::Baseline;
    const double dt = phys == Physics::Extended ? dt_cfl_ext(mat, dom, rs.kmult) : dt_cfl(mat, dom);
    const int n_steps = static_cast<int>(t_end / dt);
    StepConst k = make_const(mat, pr, dom, dt);
    k.kmult = rs.kmult;
    return RowSetup{dom, k, phys == Physics::Extended ? 1 : 0, mat.Tm, pr.Tamb,
                    single_track_ticks(pr.V, dt, dom.Lx, dom.Ly, n_steps, 0)};
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::milliseconds>(
               std::chrono::steady_clock::now() - t0).count();
}

// One run without disk output (reference run_one_row, hh_gpu_main.cpp:170-198).
RowResult run_one_row(RunSpec const& rs, double t_end, CLI const& cli) {
    RowSetup su = row_setup(rs, t_end);
    const long n_steps = static_cast<long>(su.plan.size());
    RunOptions o = run_options(cli);
    o.keep_records = false;
    DeviceRun run(su.dom, su.k, su.phys, su.Tm, std::move(su.plan), su.Tamb, o);
    const auto w0 = std::chrono::steady_clock::now();
    run.run();
    return RowResult{run.steady_W_sub(), run.peak_T(), ms_since(w0), n_steps};
}

// --ensemble: the rows of each physics as one lpbf::Ensemble, stepped by
// one launch per step for all of them.  Every value equals the row-by-row
// loop's; wall_ms is the wall time of the row's group.
std::vector<RowResult> run_rows_ensemble(std::vector<RunSpec> const& rows, CLI const& cli) {
    std::vector<RowResult> res(rows.size());
    for (int phys = 0; phys < 2; ++phys) {
        std::vector<std::size_t> idx;
        std::vector<EnsembleMember> members;
        Domain dom;
        double Tm = 0.0;
        for (std::size_t i = 0; i < rows.size(); ++i) {
            RowSetup su = row_setup(rows[i], cli.t_end);
            if (su.phys != phys) continue;
            dom = su.dom;
            Tm = su.Tm;
            members.push_back(EnsembleMember{su.k, std::move(su.plan), su.Tamb});
            idx.push_back(i);
        }
        if (idx.empty()) continue;
        EnsembleOptions o;
        const RunOptions ro = run_options(cli);
        o.bc = ro.bc;
        o.T_bc = ro.T_bc;
        o.batch = ro.batch;
        o.precision = ro.precision;
        if (ro.precision == 32) {   // the FP32 kernel's tiling; the FP64 launch picks its own
            o.depth = ro.depth;
            o.rows = ro.rows;
            o.cruise_cols = ro.cruise_cols;
        }
        Ensemble ens(dom, phys, Tm, std::move(members), o);
        const auto w0 = std::chrono::steady_clock::now();
        ens.run();
        const double wall_ms = ms_since(w0);
        for (std::size_t j = 0; j < idx.size(); ++j) {
            RunMetrics const& m = ens.metrics(static_cast<int>(j));
            res[idx[j]] = RowResult{m.steady_W_sub(), m.peak_T(), wall_ms,
                                    static_cast<long>(ens.n_steps(static_cast<int>(j)))};
        }
    }
    return res;
}

void emit_result_line(RunSpec const& rs, RowResult const& r) {
    std::printf("RESULT,%s,%s,%s,%s,%s,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%.3f,%ld\n",
                rs.run_id.c_str(), rs.cond_id.c_str(), rs.draw_id.c_str(), rs.config.c_str(),
                rs.physics.c_str(), rs.P, rs.V, rs.r0, rs.eta, rs.dsub_m, rs.kmult, r.steady_W,
                r.peak_T, r.wall_ms, r.n_steps);
    std::fflush(stdout);
}

std::vector<std::string> split_csv(std::string const& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        if (ch == ',') { out.push_back(cur); cur.clear(); }
        else if (ch != '\r') cur.push_back(ch);
    }
    out.push_back(cur);
    return out;
}

// Manifest loop (reference run_manifest, hh_gpu_main.cpp:234-290): columns
// are found by header name; one RESULT line per row.
int run_manifest(CLI const& cli) {
    std::ifstream f(cli.manifest);
    if (!f) { std::cerr << "cannot open manifest '" << cli.manifest << "'\n"; return 4; }
    std::string header;
    if (!std::getline(f, header)) { std::cerr << "empty manifest\n"; return 4; }
    const std::vector<std::string> cols = split_csv(header);
    auto col = [&](const char* name) -> int {
        for (std::size_t i = 0; i < cols.size(); ++i)
            if (cols[i] == name) return static_cast<int>(i);
        std::cerr << "manifest missing column '" << name << "'\n";
        std::exit(4);
    };
    const int i_run = col("run_id"), i_cond = col("cond_id"), i_draw = col("draw_id"),
              i_cfg = col("config"), i_P = col("P_W"), i_V = col("V_mps"), i_r0 = col("r0_m"),
              i_eta = col("eta"), i_dsub = col("dsub_m"), i_km = col("kmult"),
              i_phys = col("physics");
    const int need = 1 + std::max({i_run, i_cond, i_draw, i_cfg, i_P, i_V, i_r0, i_eta, i_dsub,
                                   i_km, i_phys});
    std::printf("# RESULT,run_id,cond_id,draw_id,config,physics,P_W,V_mps,r0_m,eta,"
                "dsub_m,kmult,steady_W_sub_um,peak_T_K,wall_ms,n_steps\n");
    std::fflush(stdout);
    std::string line;
    std::vector<RunSpec> rows;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        const std::vector<std::string> v = split_csv(line);
        if (static_cast<int>(v.size()) < need) {
            std::cerr << "manifest row " << rows.size() + 1 << " has " << v.size() << " fields\n";
            return 4;
        }
        RunSpec rs;
        rs.run_id = v[i_run]; rs.cond_id = v[i_cond]; rs.draw_id = v[i_draw];
        rs.config = v[i_cfg]; rs.physics = v[i_phys];
        rs.P = std::atof(v[i_P].c_str()); rs.V = std::atof(v[i_V].c_str());
        rs.r0 = std::atof(v[i_r0].c_str()); rs.eta = std::atof(v[i_eta].c_str());
        rs.dsub_m = std::atof(v[i_dsub].c_str()); rs.kmult = std::atof(v[i_km].c_str());
        rows.push_back(rs);
    }
    if (cli.ensemble) {
        const std::vector<RowResult> res = run_rows_ensemble(rows, cli);
        for (std::size_t i = 0; i < rows.size(); ++i) emit_result_line(rows[i], res[i]);
    } else if (cli.jobs <= 1) {
        for (RunSpec const& rs : rows) emit_result_line(rs, run_one_row(rs, cli.t_end, cli));
    } else {
        // worker threads take rows in order; each row owns a DeviceRun and its
        // stream, so small plates share the GPU; lines are printed in order
        std::vector<RowResult> res(rows.size());
        std::vector<std::exception_ptr> err(rows.size());
        std::vector<char> ready(rows.size(), 0);
        std::mutex m;
        std::condition_variable cv;
        std::atomic<std::size_t> next{0};
        std::vector<std::thread> pool;
        const int nthreads = std::min<int>(cli.jobs, static_cast<int>(rows.size()));
        for (int t = 0; t < nthreads; ++t)
            pool.emplace_back([&] {
                const bool dev_ok = hipSetDevice(cli.device) == hipSuccess;
                for (std::size_t i; (i = next.fetch_add(1)) < rows.size();) {
                    RowResult r{};
                    std::exception_ptr e;
                    try {
                        if (!dev_ok) throw std::runtime_error("hipSetDevice failed in a worker");
                        r = run_one_row(rows[i], cli.t_end, cli);
                    } catch (...) {
                        e = std::current_exception();
                    }
                    {
                        std::lock_guard<std::mutex> g(m);
                        res[i] = r;
                        err[i] = e;
                        ready[i] = 1;
                    }
                    cv.notify_all();
                }
            });
        int rc = 0;
        for (std::size_t i = 0; i < rows.size(); ++i) {
            std::unique_lock<std::mutex> g(m);
            cv.wait(g, [&] { return ready[i] != 0; });
            if (err[i]) {
                try { std::rethrow_exception(err[i]); }
                catch (std::exception const& e) { std::cerr << "lpbf_gpu: row " << i + 1 << ": " << e.what() << "\n"; }
                rc = 2;
                continue;
            }
            g.unlock();
            emit_result_line(rows[i], res[i]);
        }
        for (auto& t : pool) t.join();
        if (rc) return rc;
    }
    std::fprintf(stderr, "[manifest] completed %zu runs\n", rows.size());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const CLI cli = parse_cli(argc, argv);
    Material mat;
    Process pr;
    Domain dom;
    if (!apply_case(cli.case_id, pr)) {
        std::cerr << "unknown case '" << cli.case_id << "'\n";
        return 1;
    }
    if (cli.P_ovr > 0) pr.P = cli.P_ovr;
    if (cli.V_ovr > 0) pr.V = cli.V_ovr;
    if (cli.r0_ovr > 0) pr.r0 = cli.r0_ovr;
    if (cli.eta_ovr > 0) mat.eta = cli.eta_ovr;
    if (cli.dsub_ovr > 0) pr.dSub = cli.dsub_ovr;
    if (cli.hconv_ovr >= 0) pr.hConv = cli.hconv_ovr;

    const Physics phys = cli.physics == "ext" ? Physics::Extended : Physics::Baseline;
    const double dt = phys == Physics::Extended ? dt_cfl_ext(mat, dom, cli.kmult) : dt_cfl(mat, dom);
    const int n_steps = static_cast<int>(cli.t_end / dt);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        std::cerr << "No HIP device found.\n";
        return 3;
    }
    if (cli.device < 0 || cli.device >= ndev) {
        std::cerr << "Invalid --device=" << cli.device << " (have " << ndev << ")\n";
        return 3;
    }
    if (hipSetDevice(cli.device) != hipSuccess) return 3;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, cli.device) != hipSuccess) return 3;

    try {
        if (!cli.manifest.empty()) return run_manifest(cli);
        if (cli.quiet) {
            RunSpec rs;
            rs.run_id = cli.case_id; rs.cond_id = cli.case_id; rs.draw_id = "0";
            rs.config = "single"; rs.physics = cli.physics;
            rs.P = pr.P; rs.V = pr.V; rs.r0 = pr.r0;
            rs.eta = mat.eta; rs.dsub_m = pr.dSub; rs.kmult = cli.kmult;
            emit_result_line(rs, run_one_row(rs, cli.t_end, cli));
            return 0;
        }

        const bool write_outputs = !cli.do_bench;
        const std::string dir = !cli.out_dir.empty()
            ? cli.out_dir
            : "snapshots_case" + cli.case_id + (phys == Physics::Extended ? "_ext" : "") + "_gpu";
        std::cout << "lpbf_gpu (MI355X, 2D z-averaged, device-resident)\n"
                  << "  device    : " << cli.device << " " << prop.name << "\n"
                  << "  case      : " << cli.case_id << "  (P=" << pr.P << " W, V=" << pr.V * 1e3
                  << " mm/s, r0=" << pr.r0 * 1e6 << " um)\n"
                  << "  physics   : " << (phys == Physics::Extended ? "extended" : "baseline") << "\n"
                  << "  grid      : " << dom.nx << " x " << dom.ny << "\n"
                  << "  dt        : " << dt * 1e6 << " us   n_steps=" << n_steps << "\n"
                  << "  precision : FP" << cli.precision << "\n"
                  << "  batch     : " << cli.batch << " steps/sync\n";

        StepConst k = make_const(mat, pr, dom, dt);
        k.kmult = cli.kmult;
        const std::vector<Tick> plan =
            single_track_ticks(pr.V, dt, dom.Lx, dom.Ly, n_steps, cli.snap_every);
        RunOptions o = run_options(cli);
        o.keep_records = write_outputs;
        DeviceRun run(dom, k, phys == Physics::Extended ? 1 : 0, mat.Tm, plan, pr.Tamb, o);

        Graph<Snapshot, int> writers("SnapshotWriter");
        if (write_outputs) {
            std::filesystem::create_directories(dir);
            writers.inputs(std::make_shared<SnapshotWriter>(
                static_cast<std::size_t>(cli.n_threads < 1 ? 1 : cli.n_threads), dir));
            writers.executeGraph();
            run.on_snapshot([&](int step, std::vector<double>&& field) {
                writers.pushData(std::make_shared<Snapshot>(Snapshot{step, std::move(field)}));
            });
        }

        const auto w0 = std::chrono::steady_clock::now();
        run.run();
        const double wall_ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                                   std::chrono::steady_clock::now() - w0).count();
        const double steps_per_s = wall_ms > 0 ? n_steps * 1000.0 / wall_ms : 0.0;

        if (write_outputs) {
            writers.finishPushingData();
            writers.waitForTermination();
            std::ofstream csv(dir + "/metrics.csv");
            csv << "step,t_us,x_laser_um,peakT_K,W_um\n";
            for (StepRecord const& r : run.records()) {
                const Tick& tk = plan[static_cast<std::size_t>(r.step)];
                if (!tk.dump) continue;
                csv << r.step << ',' << tk.t * 1e6 << ',' << tk.x_laser_csv * 1e6 << ','
                    << r.stats.tmax << ',' << r.W_um << '\n';
            }
            std::ofstream meta(dir + "/meta.json");
            meta << "{\n"
                 << "  \"case\":     \"" << cli.case_id << "\",\n"
                 << "  \"backend\":  \"hip-device-resident\",\n"
                 << "  \"device\":   \"" << prop.name << "\",\n"
                 << "  \"physics\":  \"" << cli.physics << "\",\n"
                 << "  \"precision\": " << cli.precision << ",\n"
                 << "  \"kmult\":    " << cli.kmult << ",\n"
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
                 << "  \"threads\":  " << cli.n_threads << ",\n"
                 << "  \"batch\":    " << cli.batch << ",\n"
                 << "  \"mode\":     " << cli.mode << ",\n"
                 << "  \"fused\":    true,\n"
                 << "  \"pinned\":   true,\n"
                 << "  \"peak_T_K\": " << run.peak_T() << ",\n"
                 << "  \"peak_W_um\":" << run.peak_W() << ",\n"
                 << "  \"steady_W_sub_um\": " << run.steady_W_sub() << ",\n"
                 << "  \"wall_ms\":  " << wall_ms << ",\n"
                 << "  \"steps_per_s\": " << steps_per_s << "\n"
                 << "}\n";
        }
        if (cli.do_bench) {
            std::cout << "\n[--bench] device-resident run\n"
                      << "  wall      : " << wall_ms << " ms\n"
                      << "  steps/s   : " << steps_per_s << "\n"
                      << "  peak_T    : " << run.peak_T() << " K\n"
                      << "  steps     : " << n_steps << "\n";
            return 0;
        }
        std::cout << "\nDone.\n"
                  << "  peak T : " << run.peak_T() << " K\n"
                  << "  peak W : " << run.peak_W() << " um\n"
                  << "  steady W (subgrid): " << run.steady_W_sub() << " um\n"
                  << "  wall   : " << wall_ms << " ms  (" << steps_per_s << " steps/s)\n"
                  << "  output : " << dir << "\n";
    } catch (std::exception const& e) {
        std::cerr << "lpbf_gpu: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
