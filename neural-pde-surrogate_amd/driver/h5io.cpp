// HDF5 reading and writing for the euler2d driver (formats in h5io.hpp).
#include "h5io.hpp"

#include <stdexcept>

#ifdef NPDE_WITH_HDF5
#include <hdf5.h>
#endif

namespace npde_io {

#ifndef NPDE_WITH_HDF5

bool available() { return false; }
static void no_hdf5() { throw std::runtime_error("libnpde_io was compiled without HDF5 support"); }
void write_snapshot(const std::string&, const npde_grid_desc&, const std::vector<double>&, double, const RunInfo&) {
    no_hdf5();
}
InitInfo read_initial_condition(const std::string&, npde_grid_desc&, std::vector<double>&) {
    no_hdf5();
    return {};
}
void write_trajectory(const std::string&, const npde_grid_desc&, const std::vector<double>&, const std::vector<Level>&,
                      const RunInfo&, double, double) {
    no_hdf5();
}
void write_packed(const std::string&, int, int, int, const float*, const double*, const std::vector<std::string>&,
                  const int*, const std::string&, double) {
    no_hdf5();
}

#else

bool available() { return true; }

namespace {

constexpr double kGamma = 1.4;

// RAII for hid_t handles
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() {
        if (id >= 0) close(id);
    }
    H(const H&) = delete;
    H& operator=(const H&) = delete;
};

void attr(hid_t obj, const char* name, double v) {
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    H a(H5Acreate2(obj, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) throw std::runtime_error(std::string("could not write attribute ") + name);
    H5Awrite(a.id, H5T_NATIVE_DOUBLE, &v);
}

void attr(hid_t obj, const char* name, int v) {
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    H a(H5Acreate2(obj, name, H5T_NATIVE_INT, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) throw std::runtime_error(std::string("could not write attribute ") + name);
    H5Awrite(a.id, H5T_NATIVE_INT, &v);
}

void attr(hid_t obj, const char* name, const std::string& v) {  // fixed-length C string
    H ty(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(ty.id, v.empty() ? 1 : v.size());
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    H a(H5Acreate2(obj, name, ty.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) throw std::runtime_error(std::string("could not write attribute ") + name);
    H5Awrite(a.id, ty.id, v.c_str());
}

void dataset(hid_t file, const char* name, const double* data, std::vector<hsize_t> dims, hid_t dcpl = H5P_DEFAULT) {
    H sp(H5Screate_simple(int(dims.size()), dims.data(), nullptr), H5Sclose);
    H d(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
    if (d.id < 0) throw std::runtime_error(std::string("could not create dataset ") + name);
    H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
}

bool has_attr(hid_t obj, const char* name) { return H5Aexists(obj, name) > 0; }

double read_double(hid_t obj, const char* name) {
    if (!has_attr(obj, name)) throw std::runtime_error(std::string("missing attribute ") + name);
    double v = 0.0;
    H a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    H5Aread(a.id, H5T_NATIVE_DOUBLE, &v);  // converts from int or float
    return v;
}

int read_int(hid_t obj, const char* name) {
    if (!has_attr(obj, name)) throw std::runtime_error(std::string("missing attribute ") + name);
    int v = 0;
    H a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    H5Aread(a.id, H5T_NATIVE_INT, &v);
    return v;
}

std::string read_string(hid_t obj, const char* name) {  // fixed (C++ writer) or variable length (h5py)
    if (!has_attr(obj, name)) throw std::runtime_error(std::string("missing attribute ") + name);
    H a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    H ty(H5Aget_type(a.id), H5Tclose);
    std::string out;
    if (H5Tis_variable_str(ty.id) > 0) {
        H mem(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mem.id, H5T_VARIABLE);
        H5Tset_cset(mem.id, H5Tget_cset(ty.id));
        char* buf = nullptr;
        H5Aread(a.id, mem.id, &buf);
        if (buf) out = buf;
        H5free_memory(buf);
    } else {
        std::vector<char> buf(H5Tget_size(ty.id) + 1, '\0');
        H5Aread(a.id, ty.id, buf.data());
        out = buf.data();
    }
    return out;
}

std::vector<double> read_dataset(hid_t file, const char* name, size_t expected) {
    H d(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (d.id < 0) throw std::runtime_error(std::string("missing dataset ") + name);
    H sp(H5Dget_space(d.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp.id) != hssize_t(expected))
        throw std::runtime_error(std::string("dataset ") + name + " has the wrong size");
    std::vector<double> v(expected);
    H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    return v;
}

int parse_bc(const std::string& s) {
    if (s == "periodic") return NPDE_BC_PERIODIC;
    if (s == "transmissive") return NPDE_BC_TRANSMISSIVE;
    throw std::runtime_error("unknown boundary condition '" + s + "'");
}

const char* bc_name(int bc) { return bc == NPDE_BC_PERIODIC ? "periodic" : "transmissive"; }

void grid_attrs(hid_t f, const npde_grid_desc& g) {
    attr(f, "x_min", g.x_min);
    attr(f, "x_max", g.x_max);
    attr(f, "y_min", g.y_min);
    attr(f, "y_max", g.y_max);
    attr(f, "bc_x", std::string(bc_name(g.bc_x)));
    attr(f, "bc_y", std::string(bc_name(g.bc_y)));
}

}  // namespace

void write_snapshot(const std::string& path, const npde_grid_desc& g, const std::vector<double>& U, double t,
                    const RunInfo& info) {
    const size_t n = size_t(g.nx) * g.ny;
    std::vector<double> rho(n), u(n), v(n), p(n), x(g.nx), y(g.ny);
    for (size_t k = 0; k < n; ++k) {  // to_primitive (state2d.hpp:21-25)
        const double d = U[k], mx = U[n + k], my = U[2 * n + k], E = U[3 * n + k];
        const double uu = mx / d, vv = my / d;
        rho[k] = d;
        u[k] = uu;
        v[k] = vv;
        p[k] = (kGamma - 1.0) * (E - 0.5 * d * (uu * uu + vv * vv));
    }
    const double dx = (g.x_max - g.x_min) / g.nx, dy = (g.y_max - g.y_min) / g.ny;
    for (int i = 0; i < g.nx; ++i) x[i] = g.x_min + (i + 0.5) * dx;
    for (int j = 0; j < g.ny; ++j) y[j] = g.y_min + (j + 0.5) * dy;
    H f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("could not create " + path + " (does the folder exist?)");
    const std::vector<hsize_t> d2{hsize_t(g.ny), hsize_t(g.nx)};
    dataset(f.id, "rho", rho.data(), d2);
    dataset(f.id, "u", u.data(), d2);
    dataset(f.id, "v", v.data(), d2);
    dataset(f.id, "p", p.data(), d2);
    dataset(f.id, "x", x.data(), {hsize_t(g.nx)});
    dataset(f.id, "y", y.data(), {hsize_t(g.ny)});
    attr(f.id, "t", t);
    attr(f.id, "step", info.step);
    attr(f.id, "gamma", kGamma);
    attr(f.id, "cfl", info.cfl);
    attr(f.id, "nx", g.nx);
    attr(f.id, "ny", g.ny);
    grid_attrs(f.id, g);
    attr(f.id, "test", info.test);
    attr(f.id, "scheme", info.scheme);
    attr(f.id, "limiter", info.limiter);
}

InitInfo read_initial_condition(const std::string& path, npde_grid_desc& g, std::vector<double>& U) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H f(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("could not open " + path + " (does the file exist?)");
    g.nx = read_int(f.id, "nx");
    g.ny = read_int(f.id, "ny");
    g.x_min = read_double(f.id, "x_min");
    g.x_max = read_double(f.id, "x_max");
    g.y_min = read_double(f.id, "y_min");
    g.y_max = read_double(f.id, "y_max");
    g.bc_x = parse_bc(read_string(f.id, "bc_x"));
    g.bc_y = parse_bc(read_string(f.id, "bc_y"));
    if (g.nx <= 0 || g.ny <= 0) throw std::runtime_error("nx and ny must be positive");
    const size_t n = size_t(g.nx) * g.ny;
    const auto rho = read_dataset(f.id, "rho", n);
    const auto u = read_dataset(f.id, "u", n);
    const auto v = read_dataset(f.id, "v", n);
    const auto p = read_dataset(f.id, "p", n);
    InitInfo info;
    if (has_attr(f.id, "t_end")) info.t_end = read_double(f.id, "t_end");
    if (has_attr(f.id, "test")) info.test = read_string(f.id, "test");
    U.assign(4 * n, 0.0);
    for (size_t k = 0; k < n; ++k) {
        if (!(rho[k] > 0.0) || !(p[k] > 0.0))
            throw std::runtime_error("initial condition has non-positive density or pressure");
        // to_conserved (state2d.hpp:16-19)
        U[k] = rho[k];
        U[n + k] = rho[k] * u[k];
        U[2 * n + k] = rho[k] * v[k];
        U[3 * n + k] = p[k] / (kGamma - 1.0) + 0.5 * rho[k] * (u[k] * u[k] + v[k] * v[k]);
    }
    return info;
}

void write_trajectory(const std::string& path, const npde_grid_desc& g, const std::vector<double>& times,
                      const std::vector<Level>& levels, const RunInfo& info, double t_end, double t_start) {
    H f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("could not create " + path + " (does the folder exist?)");
    dataset(f.id, "t", times.data(), {hsize_t(times.size())});
    for (const Level& L : levels) {
        hsize_t dims[4] = {hsize_t(times.size()), 4, hsize_t(L.ny_c), hsize_t(L.m)};
        H sp(H5Screate_simple(4, dims, nullptr), H5Sclose);
        H pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
        hsize_t chunk[4] = {1, 4, hsize_t(L.ny_c), hsize_t(L.m)};
        H5Pset_chunk(pl.id, 4, chunk);
        H5Pset_shuffle(pl.id);
        H5Pset_deflate(pl.id, 4);
        const std::string name = "U_" + std::to_string(L.m);
        H d(H5Dcreate2(f.id, name.c_str(), H5T_NATIVE_FLOAT, sp.id, H5P_DEFAULT, pl.id, H5P_DEFAULT), H5Dclose);
        if (d.id < 0) throw std::runtime_error("could not create dataset " + name);
        H5Dwrite(d.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, L.data.data());
    }
    attr(f.id, "variables", std::string("rho, rho*u, rho*v, E"));
    attr(f.id, "t_start", t_start);
    attr(f.id, "t_end", t_end);
    attr(f.id, "frames", int(times.size()));
    attr(f.id, "steps", info.step);
    attr(f.id, "gamma", kGamma);
    attr(f.id, "cfl", info.cfl);
    attr(f.id, "solver_nx", g.nx);
    attr(f.id, "solver_ny", g.ny);
    grid_attrs(f.id, g);
    attr(f.id, "test", info.test);
    attr(f.id, "scheme", info.scheme);
    attr(f.id, "limiter", info.limiter);
}

void write_packed(const std::string& path, int N, int F, int n, const float* U, const double* t,
                  const std::vector<std::string>& family, const int* index, const std::string& split, double dt) {
    if (N <= 0 || F <= 0 || n <= 0 || int(family.size()) != N) throw std::runtime_error("bad packed dataset shape");
    H f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("could not create " + path + " (does the folder exist?)");
    // no object timestamps: the same trajectories pack into the same bytes
    H untimed(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    H5Pset_obj_track_times(untimed.id, 0);
    {
        hsize_t dims[5] = {hsize_t(N), hsize_t(F), 4, hsize_t(n), hsize_t(n)};
        hsize_t chunk[5] = {1, 1, 4, hsize_t(n), hsize_t(n)};
        H sp(H5Screate_simple(5, dims, nullptr), H5Sclose);
        H pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
        H5Pset_obj_track_times(pl.id, 0);
        H5Pset_chunk(pl.id, 5, chunk);
        H5Pset_shuffle(pl.id);
        H5Pset_deflate(pl.id, 1);
        H d(H5Dcreate2(f.id, "U", H5T_NATIVE_FLOAT, sp.id, H5P_DEFAULT, pl.id, H5P_DEFAULT), H5Dclose);
        if (d.id < 0 || H5Dwrite(d.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, U) < 0)
            throw std::runtime_error("could not write dataset U");
    }
    dataset(f.id, "t", t, {hsize_t(N), hsize_t(F)}, untimed.id);
    {
        std::vector<const char*> ptrs(N);
        for (int k = 0; k < N; ++k) ptrs[k] = family[k].c_str();
        H ty(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(ty.id, H5T_VARIABLE);
        H5Tset_cset(ty.id, H5T_CSET_UTF8);
        hsize_t dims[1] = {hsize_t(N)};
        H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
        H d(H5Dcreate2(f.id, "family", ty.id, sp.id, H5P_DEFAULT, untimed.id, H5P_DEFAULT), H5Dclose);
        if (d.id < 0 || H5Dwrite(d.id, ty.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptrs.data()) < 0)
            throw std::runtime_error("could not write dataset family");
    }
    {
        hsize_t dims[1] = {hsize_t(N)};
        H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
        H d(H5Dcreate2(f.id, "index", H5T_STD_I32LE, sp.id, H5P_DEFAULT, untimed.id, H5P_DEFAULT), H5Dclose);
        if (d.id < 0 || H5Dwrite(d.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, index) < 0)
            throw std::runtime_error("could not write dataset index");
    }
    attr(f.id, "variables", std::string("rho, rho_u, rho_v, E"));
    attr(f.id, "split", split);
    attr(f.id, "n", n);
    attr(f.id, "frames", F);
    // one frame spacing shared by every row; rows with windows of their own (NaN) leave t authoritative
    if (dt == dt) attr(f.id, "dt", dt);
}

#endif

}  // namespace npde_io
