// Error channel of the C-ABI: the reference launcher has none (a SYCL exception ends
// the process, hydro2d/kernel.cxx:331-367), so every entry point here returns a status
// and leaves a message for hydro_last_error().
#include "hydro_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>

#include <exception>
#include <new>

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

void set_exception_error(const char* fn)
{
    try {
        throw;
    } catch (const std::bad_alloc&) {
        set_error("%s: out of host memory", fn);
    } catch (const std::exception& e) {
        set_error("%s: %s", fn, e.what());
    } catch (...) {
        set_error("%s: unexpected exception", fn);
    }
}

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

#include <string.h>

#include "hydro_kernels.hpp"

extern "C" int hydro_tune_set(const char* key, int value)
{
    hydro::clear_error();
    const bool is = key != nullptr;
    if (is && strcmp(key, "rows2d") == 0 && (value == 0 || (value >= 4 && value <= 4096))) {
        hydro::set_rows_2d(value);
        return 0;
    }
    if (is && strcmp(key, "occ2d") == 0 && (value == 0 || (value >= 3 && value <= 8))) {
        hydro::set_occ_2d(value);
        return 0;
    }
    if (is && strcmp(key, "rows3d") == 0 && (value == 0 || (value >= 2 && value <= 4096))) {
        hydro::set_tile_3d(value);
        return 0;
    }
    hydro::set_error("hydro_tune_set: unknown key or value out of range (%s = %d)", key ? key : "(null)", value);
    return HYDRO_ERR_ARG;
}

extern "C" int hydro_tune_get(const char* key)
{
    if (!key) return HYDRO_ERR_ARG;
    if (strcmp(key, "rows2d") == 0) return hydro::get_rows_2d(false);
    if (strcmp(key, "occ2d") == 0) return hydro::get_occ_2d(false);
    if (strcmp(key, "rows3d") == 0) return hydro::get_tile_3d(false);
    if (strcmp(key, "rows2d_last") == 0) return hydro::get_rows_2d(true);
    if (strcmp(key, "occ2d_last") == 0) return hydro::get_occ_2d(true);
    if (strcmp(key, "rows3d_last") == 0) return hydro::get_tile_3d(true);
    return HYDRO_ERR_ARG;
}
