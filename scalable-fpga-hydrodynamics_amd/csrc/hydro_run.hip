// Device-resident step loop and diagnostics.
//
// hydro_run moves the host loop of hydro2d/main.cxx:686-815 (hydro3d/main.cxx:742-876)
// onto the device: Dt, t and the iteration count live in a StepControl block, each
// step's kernels read Dt from it (DtDx = Dt / DX computed on the device, the same IEEE
// division the host performs), and the one-wave finalise kernel that follows each step
// kernel turns the speed reduction into the next Dt, records (t, Dt), advances
// t += Dt and sets `done` when the loop condition fails.  The host can therefore enqueue
// many steps back to back (or capture them in a graph) without reading Dt between
// steps.  A step is the step kernel and the finalise kernel; periodic axes are read from
// the opposite interior plane inside the step kernel, so only a reflective X boundary
// adds a ghost-fill launch.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "../../include/hydro_hip.h"
#include "hydro_device.hpp"
#include "hydro_kernels.hpp"
#include "hydro_runtime.hpp"

using namespace hydro;

struct hydro_run {
    int dim = 0;
    int precision = 64;
    size_t nb[3] = { 0, 0, 0 };
    int bc[3] = { HYDRO_BC_PERIODIC, HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE };
    double consts[5] = {};  // C_min, gamma, gamma - 1, 1 / (gamma - 1), K
    StepControl* d_ctl = nullptr;
    double* d_traj_t = nullptr;
    void* d_traj_dt = nullptr;
    unsigned long long* d_flags = nullptr;  // IC state flags of run_restart_blast
    size_t max_it = 0;
    size_t cap_it = 0;  // iterations the trajectory record holds
    size_t enqueued = 0;
    // Optional per-launch timing of the step kernel (HIP events on the run's stream).
    bool timing = false;
    std::vector<hipEvent_t> ev;  // pairs: [2n] before, [2n+1] after launch n
    size_t ev_used = 0;          // launches recorded since the last query
};

namespace {

constexpr size_t kMaxTimed = 65536;  // launches timed between two hydro_run_kernel_time calls

// gamma, K and the derived constants of hydro2d/main.cxx:601-606 in the state precision.
void physics_constants(int precision, double C_min, double out[5])
{
    if (precision != HYDRO_FP32) {
        const double gamma = 1.4, K = 1.1;
        const double gmo = gamma - 1.0;
        out[0] = C_min;
        out[1] = gamma;
        out[2] = gmo;
        out[3] = 1.0 / gmo;
        out[4] = K;
    } else {
        const float gamma = 1.4f, K = 1.1f;
        const float gmo = gamma - 1.0f;
        out[0] = static_cast<float>(C_min);
        out[1] = gamma;
        out[2] = gmo;
        out[3] = 1.0f / gmo;
        out[4] = K;
    }
}

// The loop control at the start of a run (hydro2d/main.cxx:686-689): Dt = dt0 (rounded
// to the state precision), t = 0, it = 0, and done already when the loop condition fails.
StepControl initial_control(int dim, int precision, const double* D, const double consts[5], double dt0, double max_t,
                            size_t max_it)
{
    StepControl h;
    memset(&h, 0, sizeof(h));
    const bool f32_state = precision == HYDRO_FP32;
    h.dt = f32_state ? static_cast<double>(static_cast<float>(dt0)) : dt0;
    h.t = 0.0;
    h.max_t = max_t;
    h.it = 0;
    h.max_it = max_it;
    h.C_min = consts[0];
    for (int d = 0; d < dim; ++d) h.D[d] = f32_state ? static_cast<double>(static_cast<float>(D[d])) : D[d];
    h.done = (0.0 < max_t && 0 < max_it) ? 0u : 1u;  // hydro2d/main.cxx:689
    return h;
}

// Checks shared by hydro_run_create and hydro_run_noalloc; `what` names the caller.
bool run_arguments_ok(const char* what, int dim, const size_t* nb, const double* D, int precision, const int* bc)
{
    if ((dim != 2 && dim != 3) || !nb || !D ||
        (precision != HYDRO_FP64 && precision != HYDRO_FP32 && precision != HYDRO_FP32_FLUX)) {
        set_error("%s: bad dim/precision/pointer arguments", what);
        return false;
    }
    for (int d = 0; d < dim; ++d) {
        if (nb[d] < 1) {
            set_error("%s: nb[%d] must be >= 1", what, d);
            return false;
        }
        if (bc && bc[d] != HYDRO_BC_PERIODIC && bc[d] != HYDRO_BC_REFLECTIVE) {
            set_error("%s: unknown boundary mode %d on axis %d", what, bc[d], d);
            return false;
        }
    }
    if (!grid_fits(dim, nb)) {
        set_error("%s: grid of more than 2^40 cells", what);
        return false;
    }
    return true;
}

// Enqueue one resident step on s: the reflective X ghost fill when X is not periodic
// (periodic axes are read from the opposite interior plane inside the step kernel, and
// Y/Z walls are applied inside it), the step kernel and the finalise kernel.
int enqueue_step(int dim, const size_t* nb, const int* bc, int precision, const double* consts, StepControl* ctl,
                 double* traj_t, void* traj_dt, void* cr, void* cm, void* nr, void* nm, hipStream_t s,
                 hipEvent_t ev_a, hipEvent_t ev_b)
{
    const bool px = bc[0] == HYDRO_BC_PERIODIC;
    const int wraps = (px ? kWrapX : 0) | (bc[1] == HYDRO_BC_PERIODIC ? kWrapY : 0) |
                      (dim == 3 && bc[2] == HYDRO_BC_PERIODIC ? kWrapZ : 0);
    int rc = 0;
    if (dim == 2) {
        if (!px) rc = ghost2d_any(cr, cm, nb[0], nb[1], bc[0], HYDRO_BC_REFLECTIVE, precision, ctl, s);
        if (!rc)
            rc = step2d_resident(cr, cm, nr, nm, nb[0], nb[1], bc[1], precision, consts, ctl, &ctl->red, traj_t,
                                 traj_dt, s, ev_a, ev_b, wraps);
    } else {
        if (!px)
            rc = ghost3d_any(cr, cm, nb[0], nb[1], nb[2], bc[0], HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE, precision,
                             ctl, s);
        if (!rc)
            rc = step3d_resident(cr, cm, nr, nm, nb[0], nb[1], nb[2], bc[1], bc[2], precision, consts, ctl, &ctl->red,
                                 traj_t, traj_dt, s, ev_a, ev_b, wraps);
    }
    return rc;
}

}  // namespace

extern "C" hydro_run* hydro_run_create(int dim, const size_t* nb, const double* D, double C_min, double dt0,
                                       double max_t, size_t max_it, int precision, const int* bc)
try {
    clear_error();
    if (!run_arguments_ok("hydro_run_create", dim, nb, D, precision, bc)) return nullptr;
    // The trajectory arrays hold max_it entries of 8 bytes; a count whose size overflows
    // would give a short buffer that the steps then write past.
    if (max_it > SIZE_MAX / sizeof(double)) {
        set_error("hydro_run_create: max_it %zu is too large for its trajectory arrays", max_it);
        return nullptr;
    }
    hydro_run* r = new (std::nothrow) hydro_run();
    if (!r) {
        set_error("hydro_run_create: out of host memory");
        return nullptr;
    }
    r->dim = dim;
    r->precision = precision;
    for (int d = 0; d < dim; ++d) {
        r->nb[d] = nb[d];
        if (bc) r->bc[d] = bc[d];
    }
    r->max_it = max_it;
    r->cap_it = max_it;
    physics_constants(precision, C_min, r->consts);

    const StepControl h = initial_control(dim, precision, D, r->consts, dt0, max_t, max_it);
    const bool f32_state = precision == HYDRO_FP32;
    const size_t traj_bytes_dt = (max_it ? max_it : 1) * (f32_state ? 4 : 8);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&r->d_ctl), sizeof(StepControl));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&r->d_traj_t), (max_it ? max_it : 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&r->d_traj_dt, traj_bytes_dt);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&r->d_flags), sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(r->d_ctl, &h, sizeof(h), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("hydro_run_create: %s", hipGetErrorString(e));
        hydro_run_destroy(r);
        return nullptr;
    }
    return r;
} catch (...) {
    hydro::set_exception_error(__func__);
    return nullptr;
}

extern "C" int hydro_run_advance(hydro_run* r, void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b, size_t n_steps,
                                 void* stream)
try {
    clear_error();
    if (!r || !rhoE_a || !mom_a || !rhoE_b || !mom_b) {
        set_error("hydro_run_advance: null argument");
        return HYDRO_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (size_t n = 0; n < n_steps; ++n) {
        const bool in_a = (r->enqueued % 2) == 0;
        void* cr = in_a ? rhoE_a : rhoE_b;
        void* cm = in_a ? mom_a : mom_b;
        void* nr = in_a ? rhoE_b : rhoE_a;
        void* nm = in_a ? mom_b : mom_a;
        int rc;
        // Timed launches until the record holds kMaxTimed of them (events are reused after
        // each hydro_run_kernel_time, so the record never grows past that).
        const bool timed = r->timing && r->ev_used < kMaxTimed;
        if (timed) {
            if (r->ev.size() < 2 * (r->ev_used + 1)) {
                for (int q = 0; q < 2; ++q) {
                    hipEvent_t e = nullptr;
                    HYDRO_HIP_OR(hipEventCreate(&e), HYDRO_ERR_HIP);
                    r->ev.push_back(e);
                }
            }
        }
        hipEvent_t ev_a = timed ? r->ev[2 * r->ev_used] : nullptr;
        hipEvent_t ev_b = timed ? r->ev[2 * r->ev_used + 1] : nullptr;
        rc = enqueue_step(r->dim, r->nb, r->bc, r->precision, r->consts, r->d_ctl, r->d_traj_t, r->d_traj_dt, cr, cm,
                          nr, nm, s, ev_a, ev_b);
        if (!rc && ev_b) ++r->ev_used;
        if (rc) return rc;
        ++r->enqueued;
    }
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_run_status(hydro_run* r, size_t* it, double* t, double* dt, int* done, int* current,
                                void* stream)
try {
    clear_error();
    if (!r) {
        set_error("hydro_run_status: null run");
        return HYDRO_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    StepControl h;
    HYDRO_HIP_OR(hipMemcpyAsync(&h, r->d_ctl, sizeof(h), hipMemcpyDeviceToHost, s), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipStreamSynchronize(s), HYDRO_ERR_HIP);
    if (it) *it = static_cast<size_t>(h.it);
    if (t) *t = h.t;
    if (dt) *dt = h.dt;
    if (done) *done = h.done ? 1 : 0;
    if (current) *current = static_cast<int>(h.it % 2);
    if (h.nonfinite) {
        set_error("hydro run: Dt became non-finite at iteration %llu (NaN or zero signal speed)",
                  static_cast<unsigned long long>(h.it ? h.it - 1 : 0));
        return HYDRO_ERR_NONFINITE;
    }
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_run_trajectory(hydro_run* r, size_t from, size_t count, double* t_out, double* dt_out,
                                    void* stream)
try {
    clear_error();
    // from > max_it first, so that max_it - from cannot wrap and from + count is never formed.
    if (!r || from > r->max_it || count > r->max_it - from) {
        set_error("hydro_run_trajectory: range of %zu iterations from %zu outside the %zu recorded iterations",
                  count, from, r ? r->max_it : 0);
        return HYDRO_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (count == 0) return 0;
    if (t_out)
        HYDRO_HIP_OR(hipMemcpyAsync(t_out, r->d_traj_t + from, count * sizeof(double), hipMemcpyDeviceToHost, s),
                     HYDRO_ERR_HIP);
    if (dt_out) {
        if (r->precision != HYDRO_FP32) {
            HYDRO_HIP_OR(hipMemcpyAsync(dt_out, static_cast<double*>(r->d_traj_dt) + from, count * sizeof(double),
                                        hipMemcpyDeviceToHost, s),
                         HYDRO_ERR_HIP);
        } else {
            float* tmp = static_cast<float*>(malloc(count * sizeof(float)));
            if (!tmp) {
                set_error("hydro_run_trajectory: out of host memory");
                return HYDRO_ERR_NOMEM;
            }
            hipError_t e = hipMemcpyAsync(tmp, static_cast<float*>(r->d_traj_dt) + from, count * sizeof(float),
                                          hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                free(tmp);
                set_error("hydro_run_trajectory: %s", hipGetErrorString(e));
                return HYDRO_ERR_HIP;
            }
            for (size_t i = 0; i < count; ++i) dt_out[i] = tmp[i];
            free(tmp);
        }
    }
    HYDRO_HIP_OR(hipStreamSynchronize(s), HYDRO_ERR_HIP);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_run_set_timing(hydro_run* r, int enable)
try {
    clear_error();
    if (!r) {
        set_error("hydro_run_set_timing: null run");
        return HYDRO_ERR_ARG;
    }
    r->timing = enable != 0;
    r->ev_used = 0;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_run_kernel_time(hydro_run* r, double* total_ms, size_t* launches)
try {
    clear_error();
    if (!r || !total_ms) {
        set_error("hydro_run_kernel_time: null argument");
        return HYDRO_ERR_ARG;
    }
    double sum = 0.0;
    for (size_t n = 0; n < r->ev_used; ++n) {
        HYDRO_HIP_OR(hipEventSynchronize(r->ev[2 * n + 1]), HYDRO_ERR_HIP);
        float ms = 0.f;
        HYDRO_HIP_OR(hipEventElapsedTime(&ms, r->ev[2 * n], r->ev[2 * n + 1]), HYDRO_ERR_HIP);
        sum += ms;
    }
    *total_ms = sum;
    if (launches) *launches = r->ev_used;
    r->ev_used = 0;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" void hydro_run_destroy(hydro_run* r)
try {
    if (!r) return;
    for (hipEvent_t e : r->ev) (void)hipEventDestroy(e);
    if (r->d_ctl) (void)hipFree(r->d_ctl);
    if (r->d_traj_t) (void)hipFree(r->d_traj_t);
    if (r->d_traj_dt) (void)hipFree(r->d_traj_dt);
    if (r->d_flags) (void)hipFree(r->d_flags);
    delete r;
} catch (...) {
    hydro::set_exception_error(__func__);
}

// ---------------------------------------------------------------------------------
// Restarting a run for a new blast without allocating or waiting (the served form's
// large requests, which would otherwise create and destroy a run each: hipFree waits
// for the whole device).
// ---------------------------------------------------------------------------------
namespace {

__global__ void clear_flags_kernel(unsigned long long* flags) { *flags = 0ull; }

// The loop control of a restarted run: h with Dt = min over the IC's states present
// (flags bit 0: inside, bit 1: outside) of their C_min / speed, as the reference's
// std::min over cells starting from 1e20 (hydro2d/main.cxx:318-337); reduction slots
// cleared (one wave, lane q clears slot q).
__global__ void __launch_bounds__(kWave) start_control_kernel(StepControl* ctl, const unsigned long long* flags,
                                                              double dt_in, double dt_out, StepControl h)
{
    ctl->red.slot[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) {
        double m = 1e20;
        if (*flags & 1ull) m = (dt_in < m) ? dt_in : m;
        if (*flags & 2ull) m = (dt_out < m) ? dt_out : m;
        ctl->dt = m;
        ctl->t = h.t;
        ctl->max_t = h.max_t;
        ctl->it = h.it;
        ctl->max_it = h.max_it;
        ctl->C_min = h.C_min;
        ctl->D[0] = h.D[0];
        ctl->D[1] = h.D[1];
        ctl->D[2] = h.D[2];
        ctl->done = h.done;
        ctl->nonfinite = 0u;
    }
}

}  // namespace

namespace hydro {

size_t run_capacity(const hydro_run* r) { return r ? r->cap_it : 0; }

int run_restart_blast(hydro_run* r, int dim, const size_t* nb, int precision, double C, double radius_frac,
                      double rho_in, double rho_out, double p_in, double p_out, double max_t, size_t max_it,
                      void* rhoE, void* mom, hipStream_t s)
{
    if (!r || !rhoE || !mom || (precision != HYDRO_FP64 && precision != HYDRO_FP32_FLUX) || max_it > r->cap_it) {
        set_error("run_restart_blast: bad run, buffers, precision or max_it %zu (record of %zu)", max_it,
                  r ? r->cap_it : 0);
        return HYDRO_ERR_ARG;
    }
    // Spacings and C_min as hydro_batch_add_blast (hydro2d/main.cxx:482-484, 595-607).
    double D[3] = { 0, 0, 0 }, ms = 0;
    for (int d = 0; d < dim; ++d) {
        D[d] = 1.0 / static_cast<double>(nb[d] + 1);
        ms = (d == 0 || D[d] < ms) ? D[d] : ms;
    }
    const double C_min = C * ms;
    r->dim = dim;
    r->precision = precision;
    for (int d = 0; d < 3; ++d) {
        r->nb[d] = d < dim ? nb[d] : 0;
        r->bc[d] = d == 0 ? HYDRO_BC_PERIODIC : HYDRO_BC_REFLECTIVE;
    }
    r->max_it = max_it;
    r->enqueued = 0;
    physics_constants(precision, C_min, r->consts);
    const StepControl h = initial_control(dim, precision, D, r->consts, 0.0, max_t, max_it);
    hipLaunchKernelGGL(clear_flags_kernel, dim3(1), dim3(1), 0, s, r->d_flags);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    double dt_in = 0, dt_out = 0;
    const int rc = ic_blast_async_f64(static_cast<double*>(rhoE), static_cast<double*>(mom), dim, nb, rho_in,
                                      rho_out, p_in, p_out, 1, radius_frac, C_min, 1.4, r->d_flags, &dt_in, &dt_out,
                                      s);
    if (rc) return rc;
    hipLaunchKernelGGL(start_control_kernel, dim3(1), dim3(kWave), 0, s, r->d_ctl, r->d_flags, dt_in, dt_out, h);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

}  // namespace hydro

// ---------------------------------------------------------------------------------
// The resident loop without allocation (hydro_run_noalloc).  Its loop control (Dt, t,
// iteration, stop flag and the Dt reduction slots) and its (t, Dt) record live in one
// static device array of the library, `scratch`, instead of memory allocated per run:
//   [0, 1 KiB)                      StepControl
//   then kNoAllocMaxIt doubles      t per iteration
//   then kNoAllocMaxIt doubles      Dt per iteration (FP64 and FP32-flux state)
//   then kNoAllocMaxIt floats       Dt per iteration as the FP32 state's steps write it
// 1.25 MiB in all, held by every device the library runs on until the process ends.
// Host code never passes `scratch` itself (in host code the name is a stand-in, not a
// device address): each device's address comes from hipGetSymbolAddress, taken once per
// device with that device current.
// ---------------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocMaxIt = HYDRO_NOALLOC_MAX_IT;
constexpr size_t kCtlBytes = 1024;
constexpr size_t kTrajTOff = kCtlBytes;
constexpr size_t kTrajDtOff = kTrajTOff + kNoAllocMaxIt * sizeof(double);
constexpr size_t kTrajDt32Off = kTrajDtOff + kNoAllocMaxIt * sizeof(double);
constexpr size_t kScratchBytes = kTrajDt32Off + kNoAllocMaxIt * sizeof(float);
static_assert(sizeof(StepControl) <= kCtlBytes, "StepControl outgrew its scratch region");
static_assert(kTrajTOff % 256 == 0 && kTrajDtOff % 256 == 0 && kTrajDt32Off % 256 == 0, "scratch regions aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's scratch (taken on the first call on that
// device), or NULL with the error recorded.
unsigned char* scratch_address()
{
    const int dev = current_device();
    if (dev < 0) return nullptr;
    if (dev >= kMaxDevices) {
        set_error("hydro_run_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hydro_run_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

// Write the loop control at the start of a call, reduction slots included (one wave,
// lane q clears slot q): the scratch keeps whatever the previous call left.
__global__ void __launch_bounds__(kWave) init_control_kernel(StepControl* ctl, StepControl h)
{
    ctl->red.slot[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) {
        ctl->dt = h.dt;
        ctl->t = h.t;
        ctl->max_t = h.max_t;
        ctl->it = h.it;
        ctl->max_it = h.max_it;
        ctl->C_min = h.C_min;
        ctl->D[0] = h.D[0];
        ctl->D[1] = h.D[1];
        ctl->D[2] = h.D[2];
        ctl->done = h.done;
        ctl->nonfinite = h.nonfinite;
    }
}

// The FP32 state's Dt record (floats) widened to doubles, in the scratch.
__global__ void widen_kernel(const float* in, double* out, unsigned long long n)
{
    const unsigned long long q = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) out[q] = static_cast<double>(in[q]);
}

}  // namespace

extern "C" int hydro_run_noalloc(int dim, const size_t* nb, const double* D, double C_min, double dt0, double max_t,
                                 size_t max_it, int precision, const int* bc, void* rhoE_a, void* mom_a,
                                 void* rhoE_b, void* mom_b, size_t* it, double* t, double* dt, int* current,
                                 double* traj_t, double* traj_dt, void* stream)
{
    clear_error();
    if (!run_arguments_ok("hydro_run_noalloc", dim, nb, D, precision, bc)) return HYDRO_ERR_ARG;
    if (!rhoE_a || !mom_a || !rhoE_b || !mom_b) {
        set_error("hydro_run_noalloc: null state buffer");
        return HYDRO_ERR_ARG;
    }
    if (max_it > kNoAllocMaxIt) {
        set_error("hydro_run_noalloc: max_it %zu exceeds the %zu iterations its scratch records", max_it,
                  kNoAllocMaxIt);
        return HYDRO_ERR_ARG;
    }
    unsigned char* base = scratch_address();
    if (!base) return HYDRO_ERR_HIP;
    StepControl* ctl = reinterpret_cast<StepControl*>(base);
    double* d_traj_t = reinterpret_cast<double*>(base + kTrajTOff);
    double* d_traj_dt = reinterpret_cast<double*>(base + kTrajDtOff);
    float* d_traj_dt32 = reinterpret_cast<float*>(base + kTrajDt32Off);
    const bool f32_state = precision == HYDRO_FP32;

    int bcs[3] = { HYDRO_BC_PERIODIC, HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE };
    for (int d = 0; d < dim; ++d)
        if (bc) bcs[d] = bc[d];
    double consts[5];
    physics_constants(precision, C_min, consts);
    const StepControl h = initial_control(dim, precision, D, consts, dt0, max_t, max_it);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(init_control_kernel, dim3(1), dim3(kWave), 0, s, ctl, h);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    // max_it launches; those after the loop has ended are no-ops (hydro_run_advance).
    for (size_t n = 0; n < max_it; ++n) {
        const bool in_a = (n % 2) == 0;
        const int rc = enqueue_step(dim, nb, bcs, precision, consts, ctl, d_traj_t,
                                    f32_state ? static_cast<void*>(d_traj_dt32) : static_cast<void*>(d_traj_dt),
                                    in_a ? rhoE_a : rhoE_b, in_a ? mom_a : mom_b, in_a ? rhoE_b : rhoE_a,
                                    in_a ? mom_b : mom_a, s, nullptr, nullptr);
        if (rc) return rc;
    }
    StepControl out;
    HYDRO_HIP_OR(hipMemcpyAsync(&out, ctl, sizeof(out), hipMemcpyDeviceToHost, s), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipStreamSynchronize(s), HYDRO_ERR_HIP);
    const size_t done_it = static_cast<size_t>(out.it);
    if (done_it > 0 && traj_t)
        HYDRO_HIP_OR(hipMemcpyAsync(traj_t, d_traj_t, done_it * sizeof(double), hipMemcpyDeviceToHost, s),
                     HYDRO_ERR_HIP);
    if (done_it > 0 && traj_dt) {
        if (f32_state) {
            const unsigned blocks = static_cast<unsigned>((done_it + 255) / 256);
            hipLaunchKernelGGL(widen_kernel, dim3(blocks), dim3(256), 0, s, d_traj_dt32, d_traj_dt,
                               static_cast<unsigned long long>(done_it));
            HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
        }
        HYDRO_HIP_OR(hipMemcpyAsync(traj_dt, d_traj_dt, done_it * sizeof(double), hipMemcpyDeviceToHost, s),
                     HYDRO_ERR_HIP);
    }
    HYDRO_HIP_OR(hipStreamSynchronize(s), HYDRO_ERR_HIP);
    if (it) *it = done_it;
    if (t) *t = out.t;
    if (dt) *dt = out.dt;
    if (current) *current = static_cast<int>(out.it % 2);
    if (out.nonfinite) {
        set_error("hydro_run_noalloc: Dt became non-finite at iteration %llu (NaN or zero signal speed)",
                  static_cast<unsigned long long>(out.it ? out.it - 1 : 0));
        return HYDRO_ERR_NONFINITE;
    }
    return 0;
}

// ---------------------------------------------------------------------------------
// Total mass (and energy): sum of interior rho (E) with error-free two-sum accumulation
// (each lane keeps a (sum, error) pair; pairs are merged in a tree per workgroup), then
// a one-thread finish kernel over the workgroups' pairs.  The reference sums on the
// host in long double with Neumaier compensation (hydro2d/main.cxx:237-265); both are
// accurate far below the printed 8 significant digits.
// ---------------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void dd_add(double& hi, double& lo, double bhi, double blo)
{
    double s, e;
    two_sum(hi, bhi, s, e);
    e = e + lo + blo;
    two_sum(s, e, hi, lo);
}

// Interior cell (i, j, k) in [0, nb_x) x [0, nb_y) x [0, nb_z) lives at element
// base + i*sx + j*sy + k*sz of the cell array (2D: nb_z = 1, sz = 0); comp 0 sums rho,
// comp 1 sums E.
template <typename T>
__global__ void __launch_bounds__(256) mass_kernel(const T* __restrict__ rhoE, long long nb_x, long long nb_y,
                                                   long long nb_z, long long base, long long sx, long long sy,
                                                   long long sz, int comp, double* partial)
{
    const long long ncell = nb_x * nb_y * nb_z;
    double hi = 0.0, lo = 0.0;
    for (long long m = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; m < ncell;
         m += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long k = m % nb_z;
        const long long q = m / nb_z;
        const long long j = q % nb_y;
        const long long i = q / nb_y;
        const long long n = base + i * sx + j * sy + k * sz;
        dd_add(hi, lo, static_cast<double>(rhoE[2 * n + comp]), 0.0);
    }
    __shared__ double sh[256], sl[256];
    sh[threadIdx.x] = hi;
    sl[threadIdx.x] = lo;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            double h = sh[threadIdx.x], l = sl[threadIdx.x];
            dd_add(h, l, sh[threadIdx.x + off], sl[threadIdx.x + off]);
            sh[threadIdx.x] = h;
            sl[threadIdx.x] = l;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sh[0];
        partial[2 * blockIdx.x + 1] = sl[0];
    }
}

__global__ void mass_finish_kernel(const double* partial, int n, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double hi = 0.0, lo = 0.0;
    for (int b = 0; b < n; ++b) dd_add(hi, lo, partial[2 * b], partial[2 * b + 1]);
    out[0] = hi + lo;
}

}  // namespace

namespace {

// Compensated (double-double) device sum of component `comp` of the interior rhoE pairs.
int total_component(const void* d_rhoE, size_t NB_X, size_t NB_Y, size_t NB_Z, int precision, int comp, double* out,
                    void* stream, const char* fn)
{
    clear_error();
    if (precision == HYDRO_FP32_FLUX) precision = HYDRO_FP64;  // FP64 state
    const size_t nbv[3] = { NB_X, NB_Y, NB_Z };
    if (!d_rhoE || !out || NB_X < 1 || NB_Y < 1 || (precision != 64 && precision != 32) ||
        !grid_fits(NB_Z ? 3 : 2, nbv)) {
        set_error("%s: bad arguments", fn);
        return HYDRO_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nblk = 1024;
    double* d_partial = nullptr;
    HYDRO_HIP_OR(hipMalloc(reinterpret_cast<void**>(&d_partial), (2 * nblk + 1) * sizeof(double)), HYDRO_ERR_HIP);
    long long nz, sx, sy, sz;
    if (NB_Z) {
        nz = static_cast<long long>(NB_Z);
        sz = 1;
        sy = nz + 2;
        sx = (static_cast<long long>(NB_Y) + 2) * sy;
    } else {
        nz = 1;
        sz = 0;
        sy = 1;
        sx = static_cast<long long>(NB_Y) + 2;
    }
    const long long base = sx + sy + sz;
    hipError_t e = hipSuccess;
    if (precision == 64)
        hipLaunchKernelGGL(mass_kernel<double>, dim3(nblk), dim3(256), 0, s, static_cast<const double*>(d_rhoE),
                           static_cast<long long>(NB_X), static_cast<long long>(NB_Y), nz, base, sx, sy, sz, comp,
                           d_partial);
    else
        hipLaunchKernelGGL(mass_kernel<float>, dim3(nblk), dim3(256), 0, s, static_cast<const float*>(d_rhoE),
                           static_cast<long long>(NB_X), static_cast<long long>(NB_Y), nz, base, sx, sy, sz, comp,
                           d_partial);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mass_finish_kernel, dim3(1), dim3(64), 0, s, d_partial, nblk, d_partial + 2 * nblk);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_partial + 2 * nblk, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_partial);
    if (e != hipSuccess) {
        set_error("%s: %s", fn, hipGetErrorString(e));
        return HYDRO_ERR_HIP;
    }
    return 0;
}

}  // namespace

extern "C" int hydro_total_mass(const void* d_rhoE, size_t NB_X, size_t NB_Y, size_t NB_Z, int precision,
                                double* mass, void* stream)
try {
    return total_component(d_rhoE, NB_X, NB_Y, NB_Z, precision, 0, mass, stream, "hydro_total_mass");
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_total_energy(const void* d_rhoE, size_t NB_X, size_t NB_Y, size_t NB_Z, int precision,
                                  double* energy, void* stream)
try {
    return total_component(d_rhoE, NB_X, NB_Y, NB_Z, precision, 1, energy, stream, "hydro_total_energy");
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}
