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
