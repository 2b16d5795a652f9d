// Error channel of the C-ABI: the reference launcher has none (a SYCL exception ends
// the process, hydro2d/kernel.cxx:331-367), so every entry point here returns a status
// and leaves a message for hydro_last_error().
#include "hydro_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hydro_hip.h"
#include "hydro_kernels.hpp"

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

__global__ void clear_reduce_kernel(StepReduce* red) { red->slot[threadIdx.x] = 0ull; }

// Pool of reduction states of the plain launchers, one per stream that has stepped on
// the device; part of the library's device image, so the launchers allocate nothing
// for the first kPlainStates streams of a device.
constexpr int kPlainStates = 256;
__device__ StepReduce g_plain_red[kPlainStates];

// Reduction state of the plain launchers (launcher*, hydro_step*): one per (device,
// stream), and per thread for the per-thread default stream, whose handle every thread
// shares.  Steps on one stream run in order and can share a state; steps on different
// streams may overlap on the device and must not (their atomic maxima would mix, and one
// finalize would clear the other's slots).  A state is taken from the device's free
// list, then from g_plain_red, or allocated (sizeof(StepReduce) bytes) once a device's
// pool is used up.  It is bound to its stream until hydro_stream_release (which returns
// it to the free list, or frees an allocated one) or the end of the process.  It is
// cleared here on s, and by the finalize kernel after every step; the caller enqueues
// its step on s right after.
namespace {

struct StateKey {
    int dev;
    hipStream_t s;
    std::thread::id t;
    bool operator<(const StateKey& o) const
    {
        if (dev != o.dev) return dev < o.dev;
        if (s != o.s) return std::less<hipStream_t>()(s, o.s);
        return t < o.t;
    }
};

struct StatePool {
    StepReduce* base = nullptr;     // g_plain_red on this device
    int used = 0;                   // entries of g_plain_red handed out so far
    std::vector<StepReduce*> free;  // released entries of g_plain_red
};

std::mutex g_state_mu;
std::map<StateKey, StepReduce*> g_states;
std::map<int, StatePool> g_pools;

StateKey state_key(int dev, hipStream_t s)
{
    return StateKey{ dev, s, s == hipStreamPerThread ? std::this_thread::get_id() : std::thread::id() };
}

bool in_pool(const StatePool& pool, const StepReduce* p)
{
    return pool.base && p >= pool.base && p < pool.base + kPlainStates;
}

}  // namespace

StepReduce* plain_reduce_state(hipStream_t s)
{
    const int dev = current_device();
    if (dev < 0) return nullptr;
    const StateKey key = state_key(dev, s);
    std::lock_guard<std::mutex> lk(g_state_mu);
    const auto it = g_states.find(key);
    if (it != g_states.end()) return it->second;
    StepReduce* p = nullptr;
    bool allocated = false;
    try {
        StatePool& pool = g_pools[dev];
        if (!pool.base) {
            void* a = nullptr;
            HYDRO_HIP_OR(hipGetSymbolAddress(&a, HIP_SYMBOL(g_plain_red)), nullptr);
            pool.base = static_cast<StepReduce*>(a);
        }
        if (!pool.free.empty()) {
            p = pool.free.back();
            pool.free.pop_back();
        } else if (pool.used < kPlainStates) {
            p = pool.base + pool.used++;
        } else {
            HYDRO_HIP_OR(hipMalloc(reinterpret_cast<void**>(&p), sizeof(StepReduce)), nullptr);
            allocated = true;
        }
        // On s, so that the clear is ordered before the first step on s; by a kernel
        // rather than a memset, which a caller capturing s would record as a memset node.
        hipLaunchKernelGGL(clear_reduce_kernel, dim3(1), dim3(kRedSlots), 0, s, p);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("reduction state: clear kernel launch failed: %s", hipGetErrorString(e));
            if (allocated) (void)hipFree(p);
            else pool.free.push_back(p);
            return nullptr;
        }
        g_states.emplace(key, p);
    } catch (...) {
        if (allocated) (void)hipFree(p);
        set_error("reduction state: out of host memory");
        return nullptr;
    }
    return p;
}

}  // namespace hydro

extern "C" int hydro_stream_release(void* stream)
try {
    using namespace hydro;
    clear_error();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int dev = current_device();
    if (dev < 0) return HYDRO_ERR_HIP;
    StepReduce* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_state_mu);
        const auto it = g_states.find(state_key(dev, s));
        if (it == g_states.end()) return 0;  // no step ran on this stream
        p = it->second;
        g_states.erase(it);  // unbound: no other stream can take it before it is returned
    }
    // The stream's steps must be done with the state before another stream takes it
    // (waited for without the lock, so other threads' launches are not held up).
    HYDRO_HIP_OR(hipStreamSynchronize(s), HYDRO_ERR_HIP);
    std::lock_guard<std::mutex> lk(g_state_mu);
    StatePool& pool = g_pools[dev];
    if (in_pool(pool, p)) {
        pool.free.push_back(p);
    } else {
        HYDRO_HIP_OR(hipFree(p), HYDRO_ERR_HIP);
    }
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_stream_states(int* bound, int* pooled_free)
{
    using namespace hydro;
    const int dev = current_device();
    if (dev < 0) return HYDRO_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_state_mu);
    int n = 0;
    for (const auto& kv : g_states)
        if (kv.first.dev == dev) ++n;
    if (bound) *bound = n;
    if (pooled_free) {
        const auto it = g_pools.find(dev);
        *pooled_free = it == g_pools.end()
                           ? kPlainStates
                           : kPlainStates - it->second.used + static_cast<int>(it->second.free.size());
    }
    return 0;
}

extern "C" const char* hydro_last_error(void) { return g_last_error; }

extern "C" int hydro_version(void) { return HYDRO_HIP_ABI_VERSION; }

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
