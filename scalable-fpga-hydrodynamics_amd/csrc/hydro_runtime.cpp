// Error channel of the C-ABI: the reference launcher has none (a SYCL exception ends
// the process, hydro2d/kernel.cxx:331-367), so every entry point here returns a status
// and leaves a message for hydro_last_error().
#include "hydro_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>

#include "../../include/hydro_hip.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace hydro {

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

int current_device()
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("hipGetDevice failed: %s", hipGetErrorString(e));
        return -1;
    }
    return dev;
}

}  // namespace hydro

extern "C" const char* hydro_last_error(void) { return g_last_error; }

extern "C" int hydro_version(void) { return HYDRO_HIP_ABI_VERSION; }
