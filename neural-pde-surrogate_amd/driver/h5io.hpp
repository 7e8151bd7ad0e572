// HDF5 files of the euler2d driver: snapshots, initial conditions and trajectories, in the
// reference's layouts (io_hdf5.hpp:54-206, trajectory.hpp:53-91), so files are interchangeable:
//   snapshot / IC   rho, u, v, p (ny, nx) binary64; x (nx), y (ny); attributes t, step, gamma, cfl,
//                   nx, ny, x_min, x_max, y_min, y_max, bc_x, bc_y, test, scheme, limiter
//   trajectory      t (frames) binary64; U_<m> (frames, 4, ny/f, m) float32 chunked per frame with
//                   shuffle + deflate 4; attributes variables, t_start, t_end, frames, steps, gamma,
//                   cfl, solver_nx, solver_ny, x/y bounds, bc_x, bc_y, test, scheme, limiter
// Every function throws std::runtime_error on failure; the driver turns that into exit code 1.
#pragma once
#include <string>
#include <vector>

#include "../../include/npde.h"

namespace npde_io {

struct RunInfo {
    std::string test, scheme, limiter;
    double cfl = 0.0;
    int step = 0;
};

struct InitInfo {
    std::string test = "init";
    double t_end = -1.0;  // negative when the file gives none
};

bool available();  // false when built without HDF5

// U: conserved planes (4, ny, nx)
void write_snapshot(const std::string& path, const npde_grid_desc& g, const std::vector<double>& U, double t,
                    const RunInfo& info);
InitInfo read_initial_condition(const std::string& path, npde_grid_desc& g, std::vector<double>& U);

struct Level {
    int m, ny_c;
    std::vector<float> data;  // frames x 4 x ny_c x m
};
void write_trajectory(const std::string& path, const npde_grid_desc& g, const std::vector<double>& times,
                      const std::vector<Level>& levels, const RunInfo& info, double t_end, double t_start);

// Packed dataset file of one (split, resolution), the layout of the reference's data/pack.py:55-80:
//   U (N, F, 4, n, n) float32 chunked (1, 1, 4, n, n), shuffle + deflate 1; t (N, F) binary64;
//   family (N) variable-length strings; index (N) int32; attributes variables, split, n, frames, dt.
void write_packed(const std::string& path, int N, int F, int n, const float* U, const double* t,
                  const std::vector<std::string>& family, const int* index, const std::string& split, double dt);

}  // namespace npde_io
