// CPU oracle for the 2D Euler finite-volume step -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++17 + OpenMP and in binary64 throughout, the arithmetic of the
// reference solver so that the HIP product can be checked against it bit for bit.  It is only ever
// loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker and never as the thing that is measured or shipped.  The product library never links it.
//
// Structure (this file's own design, not the reference's): the state is stored as four interior
// planes (rho, rho*u, rho*v, E), each ny rows of nx cells, row-major.  Ghost cells are never
// materialised: a ghost index is remapped to the interior cell it would have been copied from
// (periodic wrap or transmissive clamp), which gives exactly the values the reference's ghost fill
// produces for every stencil the update uses (corner ghosts are never read by the update).
//
// Arithmetic order follows, operation for operation (compile with -ffp-contract=off):
//   state2d.hpp:16-36   to_conserved / to_primitive / sound_speed / flux_x
//   limiters.hpp:7-18   none / unlimited / minmod / van Leer
//   solver2d.hpp:17-27  limited_slope / edge
//   riemann2d.hpp:11-44 Rusanov, HLLC (Davis speeds, Toro 10.37-10.39), y-face by momentum swap
//   solver2d.hpp:67-83  compute_dt (positivity check, max wave speeds)
//   solver2d.hpp:86-103 SSP-RK2 (Heun)
//   solver2d.hpp:129-180 compute_rhs: x faces accumulated first, then y faces
//   main2d.cpp:221-246  run loop with pre-roll to t_start and exact landing on frame times
//   trajectory.hpp:31-51 block average of conserved variables, summed rows-then-columns, to float32
//   main2d.cpp:52-135   built-in initial conditions sod_x, sod_y, lax_liu3, sine2d, kh
//
// Pinning: every function here is checked in tests/ against golden vectors written by the
// reference's own code (tests/golden/) and, where the reference sources are present, against a
// build of them (oracle/_ref); see tests/test_oracle.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kGamma = 1.4;
constexpr double kGm1 = kGamma - 1.0;  // 0.3999999999999999: 1.4 is not exact in binary64

struct Cons { double d, m, n, e; };  // rho, rho*u, rho*v, E
struct Prim { double d, u, v, p; };  // rho, u, v, p

inline Prim prim_of(const Cons& c) {
    const double u = c.m / c.d;
    const double v = c.n / c.d;
    const double kin = (0.5 * c.d) * (u * u + v * v);
    return {c.d, u, v, kGm1 * (c.e - kin)};
}

inline Cons cons_of(const Prim& w) {
    const double kin = (0.5 * w.d) * (w.u * w.u + w.v * w.v);
    return {w.d, w.d * w.u, w.d * w.v, w.p / kGm1 + kin};
}

inline double sound(const Prim& w) { return std::sqrt((kGamma * w.p) / w.d); }

// std::min / std::max semantics (first argument wins ties and NaN comparisons)
inline double tmin(double a, double b) { return (b < a) ? b : a; }
inline double tmax(double a, double b) { return (a < b) ? b : a; }

inline Cons phys_flux(const Cons& c) {
    const Prim w = prim_of(c);
    return {c.m, c.m * w.u + w.p, c.n * w.u, w.u * (c.e + w.p)};
}

inline Cons swap_mom(const Cons& c) { return {c.d, c.n, c.m, c.e}; }

// F(U) + S (U* - U) for the side (c, w) with wave speed s and contact speed ss
inline Cons hllc_side(const Cons& c, const Prim& w, double s, double ss) {
    const double q = w.d * (s - w.u);
    const double f = q / (s - ss);
    const Cons st{f, f * ss, f * w.v, f * (c.e / w.d + (ss - w.u) * (ss + w.p / q))};
    const Cons F = phys_flux(c);
    return {F.d + s * (st.d - c.d), F.m + s * (st.m - c.m), F.n + s * (st.n - c.n),
            F.e + s * (st.e - c.e)};
}

inline Cons hllc(const Cons& a, const Cons& b) {
    const Prim wa = prim_of(a), wb = prim_of(b);
    const double ca = sound(wa), cb = sound(wb);
    const double sl = tmin(wa.u - ca, wb.u - cb);
    const double sr = tmax(wa.u + ca, wb.u + cb);
    const double num = ((wb.p - wa.p) + (wa.d * wa.u) * (sl - wa.u)) - (wb.d * wb.u) * (sr - wb.u);
    const double den = wa.d * (sl - wa.u) - wb.d * (sr - wb.u);
    const double ss = num / den;
    if (0.0 <= sl) return phys_flux(a);
    if (0.0 <= ss) return hllc_side(a, wa, sl, ss);
    if (0.0 <= sr) return hllc_side(b, wb, sr, ss);
    return phys_flux(b);
}

inline Cons rusanov(const Cons& a, const Cons& b) {
    const Prim wa = prim_of(a), wb = prim_of(b);
    const double s = tmax(std::fabs(wa.u) + sound(wa), std::fabs(wb.u) + sound(wb));
    const Cons fa = phys_flux(a), fb = phys_flux(b);
    const double h = 0.5 * s;
    return {0.5 * (fa.d + fb.d) - h * (b.d - a.d), 0.5 * (fa.m + fb.m) - h * (b.m - a.m),
            0.5 * (fa.n + fb.n) - h * (b.n - a.n), 0.5 * (fa.e + fb.e) - h * (b.e - a.e)};
}

enum { LIM_NONE = 0, LIM_UNLIMITED = 1, LIM_MINMOD = 2, LIM_VANLEER = 3 };
enum { FLUX_RUSANOV = 0, FLUX_HLLC = 1 };
enum { BC_TRANSMISSIVE = 0, BC_PERIODIC = 1 };

inline double limit(int kind, double a, double b) {
    switch (kind) {
        case LIM_UNLIMITED: return 0.5 * (a + b);
        case LIM_MINMOD:
            if (a * b <= 0.0) return 0.0;
            return (std::fabs(a) < std::fabs(b)) ? a : b;
        case LIM_VANLEER:
            if (a * b <= 0.0) return 0.0;
            return 2.0 * a * b / (a + b);
        default: return 0.0;
    }
}

inline Prim slope3(int kind, const Prim& m, const Prim& c, const Prim& p) {
    return {limit(kind, c.d - m.d, p.d - c.d), limit(kind, c.u - m.u, p.u - c.u),
            limit(kind, c.v - m.v, p.v - c.v), limit(kind, c.p - m.p, p.p - c.p)};
}

// face-side state: cell value moved half a cell along its slope, sgn = +1 (right/top), -1 (left/bottom)
inline Cons side_state(const Prim& w, const Prim& dw, double sgn) {
    const double h = 0.5 * sgn;
    return cons_of({w.d + h * dw.d, w.u + h * dw.u, w.v + h * dw.v, w.p + h * dw.p});
}

inline Cons face_flux(int scheme, const Cons& a, const Cons& b) {
    return scheme == FLUX_RUSANOV ? rusanov(a, b) : hllc(a, b);
}

struct Grid {
    int nx, ny;
    double x_min, x_max, y_min, y_max;
    int bc_x, bc_y;
};

inline int remap(int i, int n, int bc) {
    if (i >= 0 && i < n) return i;
    if (bc == BC_PERIODIC) return ((i % n) + n) % n;
    return i < 0 ? 0 : n - 1;
}

struct Field {  // four interior planes
    double* q[4];
    Cons at(size_t k) const { return {q[0][k], q[1][k], q[2][k], q[3][k]}; }
    void put(size_t k, const Cons& c) const { q[0][k] = c.d; q[1][k] = c.m; q[2][k] = c.n; q[3][k] = c.e; }
};

Field planes(double* U, const Grid& g) {
    const size_t n = size_t(g.nx) * g.ny;
    return {{U, U + n, U + 2 * n, U + 3 * n}};
}

int threads_or_default(int nt) {
#ifdef _OPENMP
    return nt > 0 ? nt : omp_get_max_threads();
#else
    (void)nt;
    return 1;
#endif
}

// L(U) for every interior cell (rhs of dU/dt = L), written into L (4 planes)
void rhs(const Grid& g, int scheme, int lim, const Field& U, std::vector<Prim>& W, const Field& L, int nt) {
    const int nx = g.nx, ny = g.ny;
    const double cx = 1.0 / ((g.x_max - g.x_min) / nx);
    const double cy = 1.0 / ((g.y_max - g.y_min) / ny);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) W[size_t(j) * nx + i] = prim_of(U.at(size_t(j) * nx + i));

    auto Wat = [&](int i, int j) -> const Prim& {
        return W[size_t(remap(j, ny, g.bc_y)) * nx + remap(i, nx, g.bc_x)];
    };

#pragma omp parallel num_threads(nt)
    {
        std::vector<Cons> F(nx + 1);  // x-face fluxes of one row
        std::vector<Cons> Gb(nx), Gt(nx);  // y-face fluxes below / above one row
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) {
            // x faces i = 0..nx, face i between cells i-1 and i
            for (int i = 0; i <= nx; ++i) {
                const Prim& wl = Wat(i - 1, j);
                const Prim& wr = Wat(i, j);
                const Prim dl = slope3(lim, Wat(i - 2, j), wl, wr);
                const Prim dr = slope3(lim, wl, wr, Wat(i + 1, j));
                F[i] = face_flux(scheme, side_state(wl, dl, +1.0), side_state(wr, dr, -1.0));
            }
            // y faces j (below) and j+1 (above) for every column
            for (int i = 0; i < nx; ++i) {
                const Prim& w0 = Wat(i, j - 2);
                const Prim& w1 = Wat(i, j - 1);
                const Prim& w2 = Wat(i, j);
                const Prim& w3 = Wat(i, j + 1);
                const Prim& w4 = Wat(i, j + 2);
                const Prim d1 = slope3(lim, w0, w1, w2);
                const Prim d2 = slope3(lim, w1, w2, w3);
                const Prim d3 = slope3(lim, w2, w3, w4);
                Gb[i] = swap_mom(face_flux(scheme, swap_mom(side_state(w1, d1, +1.0)),
                                           swap_mom(side_state(w2, d2, -1.0))));
                Gt[i] = swap_mom(face_flux(scheme, swap_mom(side_state(w2, d2, +1.0)),
                                           swap_mom(side_state(w3, d3, -1.0))));
            }
            for (int i = 0; i < nx; ++i) {
                const size_t k = size_t(j) * nx + i;
                const Cons& fl = F[i];
                const Cons& fr = F[i + 1];
                Cons a{0.0 + cx * fl.d, 0.0 + cx * fl.m, 0.0 + cx * fl.n, 0.0 + cx * fl.e};
                a = {a.d - cx * fr.d, a.m - cx * fr.m, a.n - cx * fr.n, a.e - cx * fr.e};
                a = {a.d + cy * Gb[i].d, a.m + cy * Gb[i].m, a.n + cy * Gb[i].n, a.e + cy * Gb[i].e};
                a = {a.d - cy * Gt[i].d, a.m - cy * Gt[i].m, a.n - cy * Gt[i].n, a.e - cy * Gt[i].e};
                L.put(k, a);
            }
        }
    }
}

struct Wavespeed {
    double sx, sy;
    int64_t bad;  // row-major index of the first non-physical cell, or -1
};

Wavespeed wavespeed(const Grid& g, const Field& U, int nt) {
    const int nx = g.nx, ny = g.ny;
    std::vector<double> psx(ny, 0.0), psy(ny, 0.0);
    std::vector<int64_t> pbad(ny, -1);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int j = 0; j < ny; ++j) {
        double sx = 0.0, sy = 0.0;
        int64_t bad = -1;
        for (int i = 0; i < nx; ++i) {
            const size_t k = size_t(j) * nx + i;
            const Prim w = prim_of(U.at(k));
            if (!(w.d > 0.0) || !(w.p > 0.0)) {
                if (bad < 0) bad = int64_t(k);
                continue;
            }
            const double a = sound(w);
            sx = tmax(sx, std::fabs(w.u) + a);
            sy = tmax(sy, std::fabs(w.v) + a);
        }
        psx[j] = sx; psy[j] = sy; pbad[j] = bad;
    }
    Wavespeed r{0.0, 0.0, -1};
    for (int j = 0; j < ny; ++j) {
        r.sx = tmax(r.sx, psx[j]);
        r.sy = tmax(r.sy, psy[j]);
        if (r.bad < 0 && pbad[j] >= 0) r.bad = pbad[j];
    }
    return r;
}

double dt_from(const Grid& g, double cfl, const Wavespeed& w) {
    const double dx = (g.x_max - g.x_min) / g.nx;
    const double dy = (g.y_max - g.y_min) / g.ny;
    return cfl / (w.sx / dx + w.sy / dy);
}

struct Workspace {
    std::vector<double> L, U1;
    std::vector<Prim> W;
    explicit Workspace(const Grid& g)
        : L(size_t(4) * g.nx * g.ny), U1(size_t(4) * g.nx * g.ny), W(size_t(g.nx) * g.ny) {}
};

void rk2(const Grid& g, int scheme, int lim, double* Ud, double dt, Workspace& ws, int nt) {
    const size_t n = size_t(g.nx) * g.ny;
    Field U = planes(Ud, g), L = planes(ws.L.data(), g), U1 = planes(ws.U1.data(), g);
    rhs(g, scheme, lim, U, ws.W, L, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long k = 0; k < (long long)n; ++k)
        for (int q = 0; q < 4; ++q) U1.q[q][k] = U.q[q][k] + dt * L.q[q][k];
    rhs(g, scheme, lim, U1, ws.W, L, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long k = 0; k < (long long)n; ++k)
        for (int q = 0; q < 4; ++q) U.q[q][k] = 0.5 * (U.q[q][k] + U1.q[q][k] + dt * L.q[q][k]);
}

void block_average(const Grid& g, const double* Ud, int m, float* out) {
    const int f = g.nx / m, nyc = g.ny / f, nxc = m;
    const double inv = 1.0 / (double(f) * f);
    const size_t n = size_t(g.nx) * g.ny;
    for (int jc = 0; jc < nyc; ++jc)
        for (int ic = 0; ic < nxc; ++ic)
            for (int q = 0; q < 4; ++q) {
                double s = 0.0;
                for (int j = jc * f; j < (jc + 1) * f; ++j)
                    for (int i = ic * f; i < (ic + 1) * f; ++i) s += Ud[q * n + size_t(j) * g.nx + i];
                out[(size_t(q) * nyc + jc) * nxc + ic] = float(s * inv);
            }
}

}  // namespace

extern "C" {

struct orc_grid {
    int nx, ny;
    double x_min, x_max, y_min, y_max;
    int bc_x, bc_y;
};

static Grid G(const orc_grid* g) {
    return {g->nx, g->ny, g->x_min, g->x_max, g->y_min, g->y_max, g->bc_x, g->bc_y};
}

// ---- point-wise known-answer helpers (arrays of n records of 4 doubles) ----
void orc_to_primitive(const double* U, double* W, long n) {
    for (long k = 0; k < n; ++k) {
        const Prim w = prim_of({U[4 * k], U[4 * k + 1], U[4 * k + 2], U[4 * k + 3]});
        W[4 * k] = w.d; W[4 * k + 1] = w.u; W[4 * k + 2] = w.v; W[4 * k + 3] = w.p;
    }
}

void orc_to_conserved(const double* W, double* U, long n) {
    for (long k = 0; k < n; ++k) {
        const Cons c = cons_of({W[4 * k], W[4 * k + 1], W[4 * k + 2], W[4 * k + 3]});
        U[4 * k] = c.d; U[4 * k + 1] = c.m; U[4 * k + 2] = c.n; U[4 * k + 3] = c.e;
    }
}

void orc_sound_speed(const double* W, double* a, long n) {
    for (long k = 0; k < n; ++k) a[k] = sound({W[4 * k], W[4 * k + 1], W[4 * k + 2], W[4 * k + 3]});
}

void orc_limiter(int kind, const double* a, const double* b, double* out, long n) {
    for (long k = 0; k < n; ++k) out[k] = limit(kind, a[k], b[k]);
}

// dir 0: x-face flux; dir 1: y-face flux (momentum swap around the x-face solver)
void orc_face_flux(int scheme, int dir, const double* UL, const double* UR, double* F, long n) {
    for (long k = 0; k < n; ++k) {
        Cons a{UL[4 * k], UL[4 * k + 1], UL[4 * k + 2], UL[4 * k + 3]};
        Cons b{UR[4 * k], UR[4 * k + 1], UR[4 * k + 2], UR[4 * k + 3]};
        Cons f = dir == 0 ? face_flux(scheme, a, b)
                          : swap_mom(face_flux(scheme, swap_mom(a), swap_mom(b)));
        F[4 * k] = f.d; F[4 * k + 1] = f.m; F[4 * k + 2] = f.n; F[4 * k + 3] = f.e;
    }
}

void orc_physical_flux(const double* U, double* F, long n) {
    for (long k = 0; k < n; ++k) {
        Cons f = phys_flux({U[4 * k], U[4 * k + 1], U[4 * k + 2], U[4 * k + 3]});
        F[4 * k] = f.d; F[4 * k + 1] = f.m; F[4 * k + 2] = f.n; F[4 * k + 3] = f.e;
    }
}

// ---- whole-grid operations; U is 4 interior planes [4][ny][nx] ----
void orc_rhs(const orc_grid* gd, int scheme, int lim, const double* U, double* L, int nthreads) {
    const Grid g = G(gd);
    std::vector<Prim> W(size_t(g.nx) * g.ny);
    Field Uf = planes(const_cast<double*>(U), g), Lf = planes(L, g);
    rhs(g, scheme, lim, Uf, W, Lf, threads_or_default(nthreads));
}

// returns 0 and sets *dt, or returns 1 and sets *bad_i, *bad_j (first non-physical cell, row-major)
int orc_compute_dt(const orc_grid* gd, double cfl, const double* U, double* dt, double* sx, double* sy,
                   int* bad_i, int* bad_j, int nthreads) {
    const Grid g = G(gd);
    Field Uf = planes(const_cast<double*>(U), g);
    const Wavespeed w = wavespeed(g, Uf, threads_or_default(nthreads));
    if (sx) *sx = w.sx;
    if (sy) *sy = w.sy;
    if (w.bad >= 0) {
        if (bad_i) *bad_i = int(w.bad % g.nx);
        if (bad_j) *bad_j = int(w.bad / g.nx);
        return 1;
    }
    *dt = dt_from(g, cfl, w);
    return 0;
}

void orc_step(const orc_grid* gd, int scheme, int lim, double* U, double dt, int nthreads) {
    const Grid g = G(gd);
    Workspace ws(g);
    rk2(g, scheme, lim, U, dt, ws, threads_or_default(nthreads));
}

void orc_block_average(const orc_grid* gd, const double* U, int m, float* out) {
    block_average(G(gd), U, m, out);
}

// Run loop of the driver: pre-roll to t_start, then nout frames landing exactly on
// t_start + (t_end - t_start) * k / nout.  Frames (block averages at each of the n_sizes sizes,
// concatenated level by level, frame-major inside each level) go to frames_out if non-null.
// dts (length max_steps) receives every dt if non-null.  Returns 0, 1 on a non-physical state
// (bad_ij set, steps/t reflect the last completed step), 2 if max_steps would be exceeded.
int orc_run(const orc_grid* gd, int scheme, int lim, double cfl, double* U, double t_start, double t_end,
            int nout, const int* sizes, int n_sizes, float* frames_out, double* frame_t,
            double* dts, long long max_steps, long long* steps_out, double* t_out, int* bad_ij,
            int nthreads) {
    const Grid g = G(gd);
    const int nt = threads_or_default(nthreads);
    Workspace ws(g);
    Field Uf = planes(U, g);
    double t = 0.0;
    long long steps = 0;
    std::vector<size_t> level_off(n_sizes > 0 ? n_sizes : 1, 0);
    size_t acc = 0;
    for (int s = 0; s < n_sizes; ++s) {
        level_off[s] = acc;
        const int f = g.nx / sizes[s];
        acc += size_t(nout + 1) * 4 * size_t(sizes[s]) * (g.ny / f);
    }
    auto output = [&](int k) {
        if (frame_t) frame_t[k] = t;
        if (!frames_out) return;
        for (int s = 0; s < n_sizes; ++s) {
            const int f = g.nx / sizes[s];
            const size_t per = size_t(4) * sizes[s] * (g.ny / f);
            block_average(g, U, sizes[s], frames_out + level_off[s] + size_t(k) * per);
        }
    };
    auto advance = [&](double target) -> int {
        while (t < target) {
            const Wavespeed w = wavespeed(g, Uf, nt);
            if (w.bad >= 0) {
                if (bad_ij) { bad_ij[0] = int(w.bad % g.nx); bad_ij[1] = int(w.bad / g.nx); }
                return 1;
            }
            if (steps >= max_steps) return 2;
            const double dt = tmin(dt_from(g, cfl, w), target - t);
            rk2(g, scheme, lim, U, dt, ws, nt);
            if (dts) dts[steps] = dt;
            t += dt;
            ++steps;
        }
        return 0;
    };
    int rc = advance(t_start);
    if (rc == 0) {
        output(0);
        for (int k = 1; k <= nout && rc == 0; ++k) {
            const double t_next = t_start + (t_end - t_start) * k / nout;
            rc = advance(t_next);
            if (rc == 0) output(k);
        }
    }
    if (steps_out) *steps_out = steps;
    if (t_out) *t_out = t;
    return rc;
}

// ---- built-in initial conditions (interior planes), grid set by the test ----
// returns 0 on success, 1 for an unknown test; gd and t_end are filled in.
int orc_setup(const char* test, int n, orc_grid* gd, double* t_end, double* U) {
    const double PI = 3.14159265358979323846;
    std::string name(test);
    Grid g{};
    if (name == "sod_x" || name == "sod_y") {
        const int m = 4;
        const bool ax = name == "sod_x";
        g = ax ? Grid{n, m, 0.0, 1.0, 0.0, double(m) / n, BC_TRANSMISSIVE, BC_TRANSMISSIVE}
               : Grid{m, n, 0.0, double(m) / n, 0.0, 1.0, BC_TRANSMISSIVE, BC_TRANSMISSIVE};
        *t_end = 0.2;
    } else if (name == "lax_liu3") {
        g = Grid{n, n, 0.0, 1.0, 0.0, 1.0, BC_TRANSMISSIVE, BC_TRANSMISSIVE};
        *t_end = 0.3;
    } else if (name == "sine2d") {
        g = Grid{n, n, 0.0, 1.0, 0.0, 1.0, BC_PERIODIC, BC_PERIODIC};
        *t_end = 1.0;
    } else if (name == "kh") {
        g = Grid{n, n, 0.0, 1.0, 0.0, 1.0, BC_PERIODIC, BC_PERIODIC};
        *t_end = 2.0;
    } else {
        return 1;
    }
    *gd = {g.nx, g.ny, g.x_min, g.x_max, g.y_min, g.y_max, g.bc_x, g.bc_y};
    if (!U) return 0;
    const double dx = (g.x_max - g.x_min) / g.nx, dy = (g.y_max - g.y_min) / g.ny;
    Field Uf = planes(U, g);
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const double x = g.x_min + (i + 0.5) * dx, y = g.y_min + (j + 0.5) * dy;
            Prim w{};
            if (name == "sod_x" || name == "sod_y") {
                const double s = name == "sod_x" ? x : y;
                w = s < 0.5 ? Prim{1.0, 0.0, 0.0, 1.0} : Prim{0.125, 0.0, 0.0, 0.1};
            } else if (name == "lax_liu3") {
                const bool right = x > 0.5, top = y > 0.5;
                if (top) w = right ? Prim{1.5, 0.0, 0.0, 1.5} : Prim{0.5323, 1.206, 0.0, 0.3};
                else     w = right ? Prim{0.5323, 0.0, 1.206, 0.3} : Prim{0.138, 1.206, 1.206, 0.029};
            } else if (name == "sine2d") {
                w = {1.0 + 0.2 * std::sin(2.0 * PI * (x + y)), 1.0, 1.0, 1.0};
            } else {  // kh
                const double r1 = 1.0, r2 = 2.0, rm = 0.5 * (r1 - r2);
                const double u1 = 0.5, u2 = -0.5, um = 0.5 * (u1 - u2);
                const double L = 0.025;
                double rho, u;
                if (y < 0.25)      { rho = r1 - rm * std::exp((y - 0.25) / L); u = u1 - um * std::exp((y - 0.25) / L); }
                else if (y < 0.5)  { rho = r2 + rm * std::exp((0.25 - y) / L); u = u2 + um * std::exp((0.25 - y) / L); }
                else if (y < 0.75) { rho = r2 + rm * std::exp((y - 0.75) / L); u = u2 + um * std::exp((y - 0.75) / L); }
                else               { rho = r1 - rm * std::exp((0.75 - y) / L); u = u1 - um * std::exp((0.75 - y) / L); }
                w = {rho, u, 0.01 * std::sin(4.0 * PI * x), 2.5};
            }
            Uf.put(size_t(j) * g.nx + i, cons_of(w));
        }
    return 0;
}

}  // extern "C"
