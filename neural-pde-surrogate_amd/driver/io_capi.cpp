// C-ABI of libnpde_io.so: the driver's HDF5 writers for callers outside C++ (the dataset packer in
// npde/dataset.py), declared in include/npde_io.h.  Errors come back as a nonzero status and a message.
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/npde_io.h"
#include "h5io.hpp"

namespace {
thread_local std::string g_err;
}

extern "C" {

int npde_io_available(void) { return npde_io::available() ? 1 : 0; }

const char* npde_io_last_error(void) { return g_err.c_str(); }

int npde_io_write_packed(const char* path, int N, int F, int n, const float* U, const double* t,
                         const char* const* family, const int* index, const char* split, double dt) {
    if (!path || !U || !t || !family || !index || !split) {
        g_err = "null argument";
        return 1;
    }
    try {
        std::vector<std::string> fam(N > 0 ? N : 0);
        for (int k = 0; k < N; ++k) fam[k] = family[k] ? family[k] : "";
        npde_io::write_packed(path, N, F, n, U, t, fam, index, split, dt);
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
    g_err.clear();
    return 0;
}

}  // extern "C"
