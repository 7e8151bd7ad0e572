// npde-euler2d: drop-in for the reference's euler2d executable, running the hot path on the GPU.
//
// Same flags and defaults (main2d.cpp:24-43), built-in tests (main2d.cpp:52-135, via
// npde_builtin_problem), --init IC files, snapshot and trajectory output modes with the same file
// layouts and names (main2d.cpp:192-218), the same stdout lines and the "Done:" throughput line that
// scripts parse (main2d.cpp:189-190, 245, 257-260), and exit code 1 with the reference's messages on
// bad arguments, bad files and non-physical states (solver2d.hpp:116-126).
//
// The time loop (main2d.cpp:221-246) runs device-resident: a trajectory run is one
// npde_run_trajectory call (frames block-averaged on the GPU as each target is reached); snapshot
// mode advances frame by frame with npde_advance_to and writes each state.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/npde.h"
#include "h5io.hpp"

namespace {

// Every flag takes one value.  The table gives the flags and their defaults (main2d.cpp:25-28); a word
// that is not a flag followed by a value, and a flag not in the table, end the run with the reference's
// messages and exit code 1 (main2d.cpp:29-41).
struct Flag {
    const char* name;
    const char* fallback;
};
constexpr Flag kFlags[] = {{"test", "lax_liu3"}, {"n", "256"},     {"scheme", "hllc"}, {"limiter", "vanleer"},
                           {"cfl", "0.5"},       {"nout", "10"},   {"out", "runs"},    {"tend", "default"},
                           {"profile", "no"},    {"init", ""},     {"traj", ""},       {"save", ""},
                           {"tstart", "0"}};

[[noreturn]] void usage_error(const std::string& msg) {
    std::cerr << msg << '\n';
    std::exit(1);
}

std::map<std::string, std::string> read_flags(int argc, char* argv[]) {
    std::map<std::string, std::string> opt;
    for (const Flag& f : kFlags) opt.emplace(f.name, f.fallback);
    for (int k = 1; k < argc; k += 2) {
        const std::string word = argv[k];
        if (word.compare(0, 2, "--") != 0 || k + 1 == argc) usage_error("Expected '--key value', got '" + word + "'");
        const auto slot = opt.find(word.substr(2));
        if (slot == opt.end()) usage_error("Unknown option " + word);
        slot->second = argv[k + 1];
    }
    return opt;
}

// --save "128,64": the trajectory's coarse sizes, in order (the reference's reading of the list,
// trajectory.hpp:104-114: empty items are an error, a trailing comma is not).
std::vector<int> save_sizes(const std::string& list) {
    std::vector<int> sizes;
    std::istringstream in(list);
    for (std::string item; std::getline(in, item, ',');) sizes.push_back(std::stoi(item));
    return sizes;
}

[[noreturn]] void die_nonphysical(const npde_bad_cell& b) {
    std::cerr << "Non-physical state in cell (" << b.i << ", " << b.j << "): rho = " << b.rho << ", p = " << b.p
              << '\n';
    std::exit(1);
}

void check(npde_solver* s, npde_status st, const npde_bad_cell& bad) {
    if (st == NPDE_OK) return;
    if (st == NPDE_ERR_NONPHYSICAL) die_nonphysical(bad);
    std::cerr << "npde error " << int(st) << ": " << npde_last_error(s) << '\n';
    std::exit(1);
}

}  // namespace

int main(int argc, char* argv[]) {
    auto a = read_flags(argc, argv);
    const std::string test = a["test"], scheme = a["scheme"], lim_name = a["limiter"];
    const int n = std::atoi(a["n"].c_str()), nout = std::atoi(a["nout"].c_str());
    const double cfl = std::atof(a["cfl"].c_str());
    if (n <= 0 || nout <= 0 || !(cfl > 0.0)) {
        std::cerr << "n, nout and cfl must be positive\n";
        return 1;
    }
    int flux;
    if (scheme == "rusanov") flux = NPDE_FLUX_RUSANOV;
    else if (scheme == "hllc") flux = NPDE_FLUX_HLLC;
    else {
        std::cerr << "Unknown scheme '" << scheme << "' (use rusanov or hllc)\n";
        return 1;
    }
    int lim;
    if (lim_name == "none") lim = NPDE_LIM_NONE;
    else if (lim_name == "minmod") lim = NPDE_LIM_MINMOD;
    else if (lim_name == "vanleer") lim = NPDE_LIM_VANLEER;
    else if (lim_name == "unlimited") lim = NPDE_LIM_UNLIMITED;
    else {
        std::cerr << "Unknown limiter '" << lim_name << "'\n";
        return 1;
    }

    npde_grid_desc g{};
    double t_end = 0.0;
    std::vector<double> U;
    std::string run_name = test;
    if (!a["init"].empty()) {
        npde_io::InitInfo ii;
        try {
            ii = npde_io::read_initial_condition(a["init"], g, U);
        } catch (const std::exception& e) {
            std::cerr << "Error reading " << a["init"] << ": " << e.what() << '\n';
            return 1;
        }
        t_end = ii.t_end;
        run_name = ii.test;
        if (t_end <= 0.0 && a["tend"] == "default") {
            std::cerr << "The initial-condition file has no t_end; give one with --tend\n";
            return 1;
        }
    } else {
        if (npde_builtin_problem(test.c_str(), n, &g, &t_end, nullptr) != NPDE_OK) {
            std::cerr << "Unknown test '" << test << "' (use sod_x, sod_y, lax_liu3, sine2d or kh)\n";
            return 1;
        }
        U.assign(size_t(4) * g.nx * g.ny, 0.0);
        npde_builtin_problem(test.c_str(), n, &g, &t_end, U.data());
    }
    if (a["tend"] != "default") t_end = std::atof(a["tend"].c_str());
    if (!(t_end > 0.0)) {
        std::cerr << "tend must be positive\n";
        return 1;
    }

    npde_solver* s = nullptr;
    if (npde_create(&g, flux, lim, cfl, &s) != NPDE_OK) {
        std::cerr << "could not create the GPU solver (is a GPU visible?)\n";
        return 1;
    }
    std::unique_ptr<npde_solver, npde_status (*)(npde_solver*)> guard(s, npde_destroy);
    npde_io::RunInfo info{run_name, scheme, lim_name, cfl, 0};
    std::printf("%s: %d x %d cells, t_end = %g, %s + %s, CFL %g, 1 GPU\n", run_name.c_str(), g.nx, g.ny, t_end,
                scheme.c_str(), lim_name.c_str(), cfl);

    const bool traj_mode = !a["traj"].empty();
    std::vector<int> sizes;
    std::string dir;
    if (traj_mode) {
        try {
            sizes = a["save"].empty() ? std::vector<int>{g.nx} : save_sizes(a["save"]);
            for (int m : sizes) {
                if (m <= 0 || g.nx % m != 0)
                    throw std::runtime_error("save size " + std::to_string(m) + " does not divide nx = " +
                                             std::to_string(g.nx));
                if (g.ny % (g.nx / m) != 0)
                    throw std::runtime_error("save size " + std::to_string(m) + " does not divide ny evenly");
            }
            if (sizes.size() > 4) throw std::runtime_error("at most 4 save sizes");
        } catch (const std::exception& e) {
            std::cerr << "Bad --save: " << e.what() << '\n';
            return 1;
        }
    } else {
        dir = a["out"] + "/" + run_name + "_" + scheme + "_" + lim_name + "_N" + std::to_string(g.nx);
        std::filesystem::create_directories(dir);
    }
    auto snap_path = [&](int k) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "/snap_%04d.h5", k);
        return dir + buf;
    };

    const double t_start = std::atof(a["tstart"].c_str());
    if (!(t_start >= 0.0 && t_start < t_end)) {
        std::cerr << "need 0 <= tstart < t_end\n";
        return 1;
    }
    npde_bad_cell bad{};
    check(s, npde_set_state(s, 0, U.data(), NPDE_LAYOUT_PLANES), bad);
    const bool profile = a["profile"] == "yes";
    if (profile) check(s, npde_set_profiling(s, 1), bad);

    double t = 0.0;
    long long steps = 0;
    double device_ms = 0.0, output_s = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [](std::chrono::steady_clock::time_point s0) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
    };
    if (traj_mode) {
        check(s, npde_run_trajectory(s, t_start, t_end, nout, sizes.data(), int(sizes.size()), &bad), bad);
        npde_last_elapsed_ms(s, &device_ms);
        npde_get_time(s, 0, &t, &steps);
        auto o0 = std::chrono::steady_clock::now();
        std::vector<double> times(nout + 1);
        check(s, npde_get_frame_times(s, 0, times.data()), bad);
        std::vector<npde_io::Level> levels;
        for (int m : sizes) {
            const int nyc = g.ny / (g.nx / m);
            npde_io::Level L{m, nyc, std::vector<float>(size_t(nout + 1) * 4 * nyc * m)};
            check(s, npde_get_trajectory(s, 0, m, L.data.data()), bad);
            levels.push_back(std::move(L));
        }
        info.step = int(steps);
        try {
            npde_io::write_trajectory(a["traj"], g, times, levels, info, t_end, t_start);
        } catch (const std::exception& e) {
            std::cerr << "Error writing trajectory: " << e.what() << '\n';
            return 1;
        }
        output_s += lap(o0);
    } else {
        auto output = [&](int k) {
            auto o0 = std::chrono::steady_clock::now();
            check(s, npde_get_state(s, 0, U.data(), NPDE_LAYOUT_PLANES), bad);
            try {
                npde_io::write_snapshot(snap_path(k), g, U, t, info);
            } catch (const std::exception& e) {
                std::cerr << "Error writing snapshot: " << e.what() << '\n';
                std::exit(1);
            }
            output_s += lap(o0);
        };
        double ms = 0.0;
        check(s, npde_advance_to(s, t_start, &bad), bad);
        npde_last_elapsed_ms(s, &ms);
        device_ms += ms;
        npde_get_time(s, 0, &t, &steps);
        output(0);
        for (int k = 1; k <= nout; ++k) {
            const double t_next = t_start + (t_end - t_start) * k / nout;
            check(s, npde_advance_to(s, t_next, &bad), bad);
            npde_last_elapsed_ms(s, &ms);
            device_ms += ms;
            npde_get_time(s, 0, &t, &steps);
            info.step = int(steps);
            output(k);
            std::printf("  snapshot %2d/%d  t = %.4f  step %lld\n", k, nout, t, steps);
        }
    }
    const double secs = lap(t0);
    const double mcells = double(g.nx) * g.ny * steps / secs / 1e6;
    std::printf("Done: %lld steps in %.2f s (%.2f million cell-updates per second). Output in %s\n", steps, secs,
                mcells, traj_mode ? a["traj"].c_str() : (dir + "/").c_str());
    if (profile) {
        // the reference's per-phase table (solver2d.hpp:43-51): here one row per kernel, from HIP events
        // around every launch of the (unbatched, graph-free) profiled run, then the launch gaps inside the
        // device loop, the output and the host's own time
        double kms[NPDE_PROF_KINDS] = {};
        long long kn[NPDE_PROF_KINDS] = {};
        check(s, npde_get_profile(s, kms, kn), bad);
        const char* names[NPDE_PROF_KINDS] = {"stage 1 (slopes, fluxes, U1)", "stage 2 (+ U^n+1, wave speeds)",
                                               "dt reduction + run state", "frames (block averages)",
                                               "wave speeds (fresh state)", "flux network"};
        const double cells = double(g.nx) * g.ny * double(steps > 0 ? steps : 1);
        const double dev = device_ms * 1e-3;
        double kern = 0.0;
        for (double v : kms) kern += v * 1e-3;
        const double gaps = dev > kern ? dev - kern : 0.0, other = secs - dev - output_s;
        auto row = [&](const char* name, double t, long long n) {
            std::printf("%-32s %9.3f %6.1f%% %12.2f", name, t, 100.0 * t / secs, 1e9 * t / cells);
            if (n >= 0) std::printf(" %9lld", n);
            std::printf("\n");
        };
        std::printf("\n%-32s %9s %7s %12s %9s\n", "phase", "time (s)", "share", "ns per cell", "launches");
        for (int k = 0; k < NPDE_PROF_KINDS; ++k)
            if (kn[k] > 0) row(names[k], kms[k] * 1e-3, kn[k]);
        row("launch gaps (device loop)", gaps, -1);
        row("output (copy + HDF5)", output_s, -1);
        row("host (setup, sync)", other > 0 ? other : 0.0, -1);
        row("total", secs, -1);
    }
    return 0;
}
