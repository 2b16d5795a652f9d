// Device-side building blocks shared by the 2D and 3D kernels (gfx950, wave64).
//
// Arithmetic follows the reference's operation order exactly (each + - * / sqrt is a
// single IEEE round-to-nearest operation; the library is compiled with
// -ffp-contract=off so nothing is fused):
//   primitives    hydro2d/kernel.cxx:5-53,   hydro3d/kernel.cxx:5-58
//   face flux     hydro2d/kernel.cxx:74-118, hydro3d/kernel.cxx:79-130
//   signal speed  hydro2d/kernel.cxx:281-313
// Sharing is what makes this cheaper than the reference while staying bit-identical:
// a face flux depends only on the two cell states in a fixed argument order, and the
// reference evaluates each interior face twice with identical arguments (as fx2 of
// cell i-1 and fx1 of cell i, kernel.cxx:186-210), so it is evaluated once here.  The
// acoustic impedance sqrt(gamma*p*rho) of a cell is likewise computed once and reused
// by all its faces (a wall mirror keeps rho and p, hence the impedance).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hydro {

constexpr int kWave = 64;
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950 LDS per CU

// Sign bit of the (high word of the) floating-point type.
template <typename T>
constexpr unsigned int kSignMask = 0x80000000u;

// Primitive state of one cell in the frame of a face: density, normal velocity,
// tangential velocities, pressure, impedance a = sqrt((gamma*p)*rho).
template <typename T>
struct Prim2 {
    T r, un, ut, p, a;
};

template <typename T>
struct Prim3 {
    T r, un, ut1, ut2, p, a;
};

template <typename T>
__device__ __forceinline__ T rmax(T a, T b)
{
    return (a < b) ? b : a;  // std::max(a, b), hydro2d/kernel.hpp:36
}

// Math policies for the reciprocals and square roots of the scheme.
//   IeeeMath: correctly rounded ONE / x and sqrt(x), exactly as the reference computes
//             them (FP64, and FP32 for the USE_FLOAT build).
//   FastMath: the hardware approximations (v_rcp_f32, v_sqrt_f32, about 1 ulp), for
//             the FP32-flux form whose results are compared to FP64 with a tolerance.
struct IeeeMath {
    template <typename T>
    static __device__ __forceinline__ T recip(T x)
    {
        return T(1) / x;
    }
    template <typename T>
    static __device__ __forceinline__ T root(T x)
    {
        return sqrt(x);
    }
};

struct FastMath {
    static __device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float root(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// Conserved -> primitive, hydro2d/kernel.cxx:5-29; plus the impedance of compute_fluxes
// (kernel.cxx:92-93), evaluated as sqrt((gamma*p)*rho).
template <typename T, typename M = IeeeMath>
__device__ __forceinline__ void cons_to_prim2(T rho, T mx, T my, T E, T gmo, T gamma, T& u, T& v, T& p, T& a)
{
    const T inv = M::recip(rho);
    u = mx * inv;
    v = my * inv;
    const T tmp = u * u + v * v;
    const T e = E * inv - T(0.5) * tmp;
    p = gmo * rho * e;
    a = M::root(gamma * p * rho);
}

template <typename T, typename M = IeeeMath>
__device__ __forceinline__ void cons_to_prim3(T rho, T mx, T my, T mz, T E, T gmo, T gamma, T& u, T& v, T& w, T& p,
                                              T& a)
{
    const T inv = M::recip(rho);
    u = mx * inv;
    v = my * inv;
    w = mz * inv;
    const T tmp = u * u + v * v + w * w;
    const T e = E * inv - T(0.5) * tmp;
    p = gmo * rho * e;
    a = M::root(gamma * p * rho);
}

// Face flux from precomputed impedances, hydro2d/kernel.cxx:95-117.
// F = (mass, normal momentum, tangential momentum, energy) in the face frame.
template <typename T, typename M = IeeeMath>
__device__ __forceinline__ void face_flux2(const Prim2<T>& L, const Prim2<T>& R, T K, T divgamma, T& F0, T& F1,
                                           T& F2, T& F3)
{
    const T a = K * rmax(L.a, R.a);
    const T diva = M::recip(a);
    const T us = T(0.5) * ((L.un + R.un) - diva * (R.p - L.p));
    const T ps = T(0.5) * ((L.p + R.p) - a * (R.un - L.un));
    const bool up = us >= T(0);
    const T r = up ? L.r : R.r;
    const T un = up ? L.un : R.un;
    const T ut = up ? L.ut : R.ut;
    const T p = up ? L.p : R.p;
    const T Un = un * r;
    const T Ut = ut * r;
    const T UE = p * divgamma + T(0.5) * r * (un * un + ut * ut);
    F0 = r * us + T(0);
    F1 = Un * us + ps;
    F2 = Ut * us + T(0);
    F3 = UE * us + ps * us;
}

// hydro3d/kernel.cxx:101-129.
template <typename T, typename M = IeeeMath>
__device__ __forceinline__ void face_flux3(const Prim3<T>& L, const Prim3<T>& R, T K, T divgamma, T& F0, T& F1,
                                           T& F2, T& F3, T& F4)
{
    const T a = K * rmax(L.a, R.a);
    const T diva = M::recip(a);
    const T us = T(0.5) * ((L.un + R.un) - diva * (R.p - L.p));
    const T ps = T(0.5) * ((L.p + R.p) - a * (R.un - L.un));
    const bool up = us >= T(0);
    const T r = up ? L.r : R.r;
    const T un = up ? L.un : R.un;
    const T ut1 = up ? L.ut1 : R.ut1;
    const T ut2 = up ? L.ut2 : R.ut2;
    const T p = up ? L.p : R.p;
    const T Un = un * r;
    const T Ut1 = ut1 * r;
    const T Ut2 = ut2 * r;
    const T eint = p * divgamma;
    const T ekin = T(0.5) * r * (un * un + ut1 * ut1 + ut2 * ut2);
    const T UE = eint + ekin;
    F0 = r * us + T(0);
    F1 = Un * us + ps;
    F2 = Ut1 * us + T(0);
    F3 = Ut2 * us + T(0);
    F4 = UE * us + ps * us;
}

// Signal speed c_s + |u| of a conserved state, hydro2d/kernel.cxx:287-309.
template <typename T, typename M = IeeeMath>
__device__ __forceinline__ T signal_speed2(T rho, T mx, T my, T E, T gmo, T gamma)
{
    const T inv = M::recip(rho);
    const T u = mx * inv;
    const T v = my * inv;
    const T tmp = u * u + v * v;
    const T e = E * inv - T(0.5) * tmp;
    const T p = gmo * rho * e;
    const T cs = M::root(gamma * p * inv);
    const T un = M::root(tmp);
    return cs + un;
}

template <typename T, typename M = IeeeMath>
__device__ __forceinline__ T signal_speed3(T rho, T mx, T my, T mz, T E, T gmo, T gamma)
{
    const T inv = M::recip(rho);
    const T u = mx * inv;
    const T v = my * inv;
    const T w = mz * inv;
    const T tmp = u * u + v * v + w * w;
    const T e = E * inv - T(0.5) * tmp;
    const T p = gmo * rho * e;
    const T cs = M::root(gamma * p * inv);
    const T un = M::root(tmp);
    return cs + un;
}

// Order key of a signal speed for the max-reduction.  Speeds are >= 0 or NaN; the sign
// bit is cleared so that -0 orders as +0, and a NaN orders above +inf so that it
// reaches Dt (where the host reports it) instead of being dropped.
__device__ __forceinline__ uint64_t speed_key(double s)
{
    return static_cast<uint64_t>(__double_as_longlong(s)) & 0x7fffffffffffffffull;
}

__device__ __forceinline__ double key_speed(uint64_t k) { return __longlong_as_double(static_cast<long long>(k)); }

__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a < b ? b : a; }

// x with its sign bit xor-ed with `mask` (0 or the sign bit of the type's high word).
// Negation is exact, so this yields -x or x bit for bit.
__device__ __forceinline__ double xor_sign(double x, unsigned int mask)
{
    return __hiloint2double(__double2hiint(x) ^ static_cast<int>(mask), __double2loint(x));
}
__device__ __forceinline__ float xor_sign(float x, unsigned int mask)
{
    return __int_as_float(__float_as_int(x) ^ static_cast<int>(mask));
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(v), off, kWave));
        v = umax64(v, o);
    }
    return v;
}

// Lane l receives x from lane l+1 (lane 63 receives its own value).
template <typename T>
__device__ __forceinline__ T from_next_lane(T x)
{
    return __shfl_down(x, 1, kWave);
}

// Lane l receives x from lane l-1 (lane 0 receives its own value).
template <typename T>
__device__ __forceinline__ T from_prev_lane(T x)
{
    return __shfl_up(x, 1, kWave);
}

// Per-step reduction state of one solve: running max of the speed keys, spread over
// kRedSlots words (workgroup b folds into slot b % kRedSlots) so that the atomics of
// many workgroups do not serialise on one address.  The step kernel only performs
// atomic max operations; the finalise kernel that follows it in the stream (the kernel
// boundary orders and publishes the atomics) turns the max into Dt and clears the slots.
constexpr int kRedSlots = kWave;

struct StepReduce {
    unsigned long long slot[kRedSlots];
};

// Device-resident step-loop control of one solve (the host loop of
// hydro2d/main.cxx:686-815 moved onto the device).  dt is the Dt the next step uses;
// after each step the finalise kernel stores traj_t[it] = t, traj_dt[it] = Dt_new, then
// it += 1, t += Dt_new (main.cxx:813-814) and sets done once !(t < max_t && it < max_it).
// Steps launched after done are no-ops.
struct StepControl {
    double dt;
    double t;
    double max_t;
    unsigned long long it;
    unsigned long long max_it;
    double C_min;
    double D[3];  // DX, DY, DZ
    unsigned int done;
    unsigned int nonfinite;  // set when a Dt came out NaN or inf
    StepReduce red;
};

// Fold every lane's key into the workgroup maximum and that into the workgroup's slot.
// Every thread of the workgroup must call it.
__device__ __forceinline__ void block_max_to_slot(StepReduce* red, uint64_t key)
{
    __shared__ unsigned long long s_wave[16];
    const uint64_t w = wave_max_u64(key);
    const int nw = static_cast<int>(blockDim.x) / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) s_wave[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        for (int q = 0; q < nw; ++q) m = umax64(m, s_wave[q]);
        if (m != 0) atomicMax(&red->slot[blockIdx.x % kRedSlots], static_cast<unsigned long long>(m));
    }
}

// Finalise one step (one wave): Dt_next = C_min / max_speed (hydro2d/kernel.cxx:318),
// clear the slots, and either store Dt_next or advance the resident loop control.
template <typename T, bool RESIDENT>
__global__ void __launch_bounds__(kWave) finalize_step_kernel(StepReduce* red, T C_min, StepControl* ctl, T* Dt_next,
                                                              double* traj_t, T* traj_dt)
{
    if constexpr (RESIDENT) {
        if (ctl->done) return;
    }
    const int lane = threadIdx.x;
    uint64_t k = red->slot[lane];
    red->slot[lane] = 0ull;
    k = wave_max_u64(k);
    if (lane != 0) return;
    const T dtn = C_min / static_cast<T>(key_speed(k));
    if constexpr (RESIDENT) {
        const unsigned long long it = ctl->it;
        if (traj_t) traj_t[it] = ctl->t;
        if (traj_dt) traj_dt[it] = dtn;
        const double t_new = ctl->t + static_cast<double>(dtn);  // hydro2d/main.cxx:813-814
        ctl->dt = static_cast<double>(dtn);
        ctl->t = t_new;
        ctl->it = it + 1;
        if (!isfinite(static_cast<double>(dtn))) {
            ctl->nonfinite = 1u;
            ctl->done = 1u;
        }
        if (!(t_new < ctl->max_t && it + 1 < ctl->max_it)) ctl->done = 1u;
    } else {
        *Dt_next = dtn;
    }
}

}  // namespace hydro
