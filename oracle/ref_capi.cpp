// C entry points over the reference solver's own headers, compiled from the sources where they lie
// (oracle/Makefile, target _ref/libref2d.so).  TEST INFRASTRUCTURE ONLY: used in this container to
// write golden vectors (tests/golden/make_golden.py) and to check the oracle, and on any machine
// where the prebuilt library is present as the "reference" CPU baseline in bench.py.
//
// The functions mirror the oracle's orc_* signatures (state = 4 interior planes [4][ny][nx]) and
// convert to and from the reference's ghosted array-of-structs layout (grid2d.hpp:12-28).
#include "solver2d.hpp"  // reference: solver/src/solver2d.hpp (+ grid2d, riemann2d, limiters, state2d)

#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <vector>

using namespace e2d;

namespace {

FluxFn pick_flux(int scheme) { return scheme == 0 ? rusanov_flux : hllc_flux; }

Limiter pick_lim(int kind) {
    switch (kind) {
        case 1: return unlimited;
        case 2: return minmod;
        case 3: return van_leer;
        default: return no_slope;
    }
}

}  // namespace

extern "C" {

struct ref_grid {
    int nx, ny;
    double x_min, x_max, y_min, y_max;
    int bc_x, bc_y;
};

static Grid mk(const ref_grid* g) {
    return make_grid(g->nx, g->ny, g->x_min, g->x_max, g->y_min, g->y_max,
                     g->bc_x ? BC::Periodic : BC::Transmissive, g->bc_y ? BC::Periodic : BC::Transmissive);
}

static void load(const Grid& g, const double* P, std::vector<Conserved>& U) {
    const size_t n = size_t(g.nx) * g.ny;
    U.assign(g.size(), Conserved{});
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const size_t k = size_t(j) * g.nx + i;
            U[g.idx(i, j)] = {P[k], P[n + k], P[2 * n + k], P[3 * n + k]};
        }
}

static void store(const Grid& g, const std::vector<Conserved>& U, double* P) {
    const size_t n = size_t(g.nx) * g.ny;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const size_t k = size_t(j) * g.nx + i;
            const Conserved& c = U[g.idx(i, j)];
            P[k] = c.rho; P[n + k] = c.mx; P[2 * n + k] = c.my; P[3 * n + k] = c.E;
        }
}

// first non-physical cell in row-major order (the scan of Solver::report_bad_cell), or -1
static int64_t first_bad(const Grid& g, const std::vector<Conserved>& U) {
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            Primitive W = to_primitive(U[g.idx(i, j)]);
            if (!(W.rho > 0.0) || !(W.p > 0.0)) return int64_t(j) * g.nx + i;
        }
    return -1;
}

void ref_to_primitive(const double* U, double* W, long n) {
    for (long k = 0; k < n; ++k) {
        Primitive w = to_primitive({U[4 * k], U[4 * k + 1], U[4 * k + 2], U[4 * k + 3]});
        W[4 * k] = w.rho; W[4 * k + 1] = w.u; W[4 * k + 2] = w.v; W[4 * k + 3] = w.p;
    }
}

void ref_to_conserved(const double* W, double* U, long n) {
    for (long k = 0; k < n; ++k) {
        Conserved c = to_conserved({W[4 * k], W[4 * k + 1], W[4 * k + 2], W[4 * k + 3]});
        U[4 * k] = c.rho; U[4 * k + 1] = c.mx; U[4 * k + 2] = c.my; U[4 * k + 3] = c.E;
    }
}

void ref_sound_speed(const double* W, double* a, long n) {
    for (long k = 0; k < n; ++k) a[k] = sound_speed({W[4 * k], W[4 * k + 1], W[4 * k + 2], W[4 * k + 3]});
}

void ref_limiter(int kind, const double* a, const double* b, double* out, long n) {
    Limiter lim = pick_lim(kind);
    for (long k = 0; k < n; ++k) out[k] = lim(a[k], b[k]);
}

void ref_face_flux(int scheme, int dir, const double* UL, const double* UR, double* F, long n) {
    FluxFn fl = pick_flux(scheme);
    for (long k = 0; k < n; ++k) {
        Conserved a{UL[4 * k], UL[4 * k + 1], UL[4 * k + 2], UL[4 * k + 3]};
        Conserved b{UR[4 * k], UR[4 * k + 1], UR[4 * k + 2], UR[4 * k + 3]};
        Conserved f = dir == 0 ? fl(a, b) : flux_y_face(fl, a, b);
        F[4 * k] = f.rho; F[4 * k + 1] = f.mx; F[4 * k + 2] = f.my; F[4 * k + 3] = f.E;
    }
}

void ref_physical_flux(const double* U, double* F, long n) {
    for (long k = 0; k < n; ++k) {
        Conserved f = flux_x({U[4 * k], U[4 * k + 1], U[4 * k + 2], U[4 * k + 3]});
        F[4 * k] = f.rho; F[4 * k + 1] = f.mx; F[4 * k + 2] = f.my; F[4 * k + 3] = f.E;
    }
}

int ref_compute_dt(const ref_grid* gd, double cfl, const double* P, double* dt, int* bad_i, int* bad_j) {
    Grid g = mk(gd);
    std::vector<Conserved> U;
    load(g, P, U);
    const int64_t b = first_bad(g, U);
    if (b >= 0) {
        if (bad_i) *bad_i = int(b % g.nx);
        if (bad_j) *bad_j = int(b / g.nx);
        return 1;
    }
    Solver s(g, hllc_flux, van_leer, cfl);
    *dt = s.compute_dt(U);
    return 0;
}

void ref_step(const ref_grid* gd, int scheme, int lim, double* P, double dt) {
    Grid g = mk(gd);
    std::vector<Conserved> U;
    load(g, P, U);
    Solver s(g, pick_flux(scheme), pick_lim(lim), 0.5);
    s.step(U, dt);
    store(g, U, P);
}

// The driver's run loop (main2d.cpp:224-246) on the reference Solver, with the trajectory levels of
// trajectory.hpp:31-51 computed by the reference's own block average (TrajectoryWriter is HDF5-bound,
// so its loop is exercised through the CLI build instead; here the frames are returned as states).
// states_out (optional) receives the full interior state at every frame: (nout+1) x 4 x ny x nx.
int ref_run(const ref_grid* gd, int scheme, int lim, double cfl, double* P, double t_start, double t_end,
            int nout, double* states_out, double* frame_t, double* dts, long long max_steps,
            long long* steps_out, double* t_out, int* bad_ij) {
    Grid g = mk(gd);
    std::vector<Conserved> U;
    load(g, P, U);
    Solver solver(g, pick_flux(scheme), pick_lim(lim), cfl);
    const size_t per = size_t(4) * g.nx * g.ny;
    double t = 0.0;
    long long step = 0;
    int rc = 0;
    auto advance = [&](double target) {
        while (t < target) {
            const int64_t b = first_bad(g, U);
            if (b >= 0) {
                if (bad_ij) { bad_ij[0] = int(b % g.nx); bad_ij[1] = int(b / g.nx); }
                return 1;
            }
            if (step >= max_steps) return 2;
            double dt = std::min(solver.compute_dt(U), target - t);
            solver.step(U, dt);
            if (dts) dts[step] = dt;
            t += dt;
            ++step;
        }
        return 0;
    };
    auto output = [&](int k) {
        if (frame_t) frame_t[k] = t;
        if (states_out) store(g, U, states_out + size_t(k) * per);
    };
    rc = advance(t_start);
    if (rc == 0) {
        output(0);
        for (int k = 1; k <= nout && rc == 0; ++k) {
            const double t_next = t_start + (t_end - t_start) * k / nout;
            rc = advance(t_next);
            if (rc == 0) output(k);
        }
    }
    store(g, U, P);
    if (steps_out) *steps_out = step;
    if (t_out) *t_out = t;
    return rc;
}

}  // extern "C"

// OpenMP threads of the following calls from this thread (the reference parallelises every loop with
// OpenMP and uses all cores by default); returns the previous setting.
extern "C" int ref_set_threads(int n) {
    const int old = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return old;
}
