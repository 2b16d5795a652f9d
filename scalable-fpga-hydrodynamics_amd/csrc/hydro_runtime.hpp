// Host-side helpers shared by the HIP translation units: error reporting through
// hydro_last_error(), per-device symbol lookups and launch validation.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace hydro {

// Record a printf-style message as this thread's last error.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
// Record the exception being handled (host allocation failures in the batch, server and
// run APIs) as this thread's last error; for the catch clauses that keep C++
// exceptions from crossing the C ABI.
void set_exception_error(const char* fn);

// Current device index, or -1 with the error recorded.
int current_device();

// Largest grid, ghosts included, that an entry point accepts: 2^40 cells, more than any
// device's memory holds, so that every cell index the kernels form (at most three values
// per cell) stays far inside 64 bits whatever sizes a caller passes.
constexpr unsigned long long kMaxGridCells = 1ull << 40;

// Whether interior sizes nb[0..dim) give at most kMaxGridCells cells with one ghost layer
// on every side (sizes below 1 are left to the caller's own checks).
inline bool grid_fits(int dim, const size_t* nb)
{
    unsigned long long cells = 1;
    for (int d = 0; d < dim; ++d) {
        if (nb[d] > kMaxGridCells) return false;
        const unsigned long long n = static_cast<unsigned long long>(nb[d]) + 2;
        if (cells > kMaxGridCells / n) return false;
        cells *= n;
    }
    return true;
}

}  // namespace hydro

#define HYDRO_HIP_OR(expr, retval)                                                                \
    do {                                                                                          \
        hipError_t hydro_e_ = (expr);                                                             \
        if (hydro_e_ != hipSuccess) {                                                             \
            hydro::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(hydro_e_), __FILE__, \
                             __LINE__);                                                           \
            return retval;                                                                        \
        }                                                                                         \
    } while (0)
