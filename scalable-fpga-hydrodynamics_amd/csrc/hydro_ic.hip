// Blast initial condition and initial Dt on the device.
//
// Reference: set_init_conditions_diffusion (hydro2d/main.cxx:280-338,
// hydro3d/main.cxx:296-363), run on the host per subdomain and followed by a full
// host->device copy (hydro2d/main.cxx:650-652).  Here every interior cell is written on
// the device and the initial Dt = min over cells of C_min / (c_s + |u|) is reduced on
// the device (hydro2d/main.cxx:316-335 and the MIN Allreduce at :660).
//
// ic_mode 0 is the reference IC: centre (NB+1)/2 per axis in 1-based global cell
// coordinates, radius = min(centres)/3, a cell is inside when its double-precision
// Euclidean distance is <= radius.  ic_mode 1 is the parameterised blast used by the
// larger and batched workloads: radius = radius_frac * min(NB + 1).
#include <hip/hip_runtime.h>

#include "../../include/hydro_hip.h"
#include "hydro_device.hpp"
#include "hydro_runtime.hpp"

namespace hydro {
namespace {

template <typename T>
__global__ void __launch_bounds__(256) ic_kernel(T* __restrict__ rhoE, T* __restrict__ mom, int dim, long long nb_x,
                                                 long long nb_y, long long nb_z, long long base_i, long long base_j,
                                                 long long base_k, double cx, double cy, double cz, double radius,
                                                 T rho_in, T rho_out, T E_in, T E_out, T dt_in, T dt_out,
                                                 unsigned long long* flags)
{
    const long long nz = dim == 3 ? nb_z : 1;
    const long long ncell = nb_x * nb_y * nz;
    const long long ys = nb_y + 2, zs = nb_z + 2;
    unsigned seen = 0;  // bit 0: an inside cell, bit 1: an outside cell
    for (long long m = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; m < ncell;
         m += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long ck = m % nz;
        const long long q = m / nz;
        const long long cj = q % nb_y;
        const long long ci = q / nb_y;
        const double di = static_cast<double>(base_i + ci + 1) - cx;
        const double dj = static_cast<double>(base_j + cj + 1) - cy;
        double d2 = di * di + dj * dj;
        if (dim == 3) {
            const double dk = static_cast<double>(base_k + ck + 1) - cz;
            d2 = d2 + dk * dk;
        }
        const bool in = sqrt(d2) <= radius;
        seen |= in ? 1u : 2u;
        const T r = in ? rho_in : rho_out;
        const T E = in ? E_in : E_out;
        if (dim == 2) {
            const long long n = (ci + 1) * ys + (cj + 1);
            rhoE[2 * n] = r;
            rhoE[2 * n + 1] = E;
            mom[2 * n] = T(0);
            mom[2 * n + 1] = T(0);
        } else {
            const long long n = ((ci + 1) * ys + (cj + 1)) * zs + (ck + 1);
            rhoE[2 * n] = r;
            rhoE[2 * n + 1] = E;
            mom[3 * n] = T(0);
            mom[3 * n + 1] = T(0);
            mom[3 * n + 2] = T(0);
        }
    }
    // Which of the two states occur (every cell of a kind has the same Dt).
    unsigned long long s = seen;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s |= __shfl_xor(s, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && s) atomicOr(flags, s);
}

}  // namespace
}  // namespace hydro

using namespace hydro;

namespace {

// Host-side parameters of the blast: centre, radius, E = p / (gamma - 1) of the two
// states and their Dt (hydro2d/main.cxx:288-333).
template <typename T>
struct IcSetup {
    double c[3];
    double radius;
    T E_in, E_out, dt_in, dt_out;
};

template <typename T>
IcSetup<T> ic_setup(int dim, const size_t* NB, T rho_in, T rho_out, T p_in, T p_out, int ic_mode, double radius_frac,
                    T C_min, T gamma)
{
    IcSetup<T> q;
    const T gmo = gamma - T(1);
    q.c[0] = q.c[1] = q.c[2] = 0;
    for (int d = 0; d < dim; ++d) q.c[d] = static_cast<double>(NB[d] + 1) / 2.0;
    if (ic_mode == 0) {
        double m = q.c[0];  // std::min({ centres }) / 3, hydro2d/main.cxx:292
        for (int d = 1; d < dim; ++d)
            if (q.c[d] < m) m = q.c[d];
        q.radius = m / 3;
    } else {
        size_t n = NB[0];
        for (int d = 1; d < dim; ++d)
            if (NB[d] < n) n = NB[d];
        q.radius = radius_frac * static_cast<double>(n + 1);
    }
    // E = p / (gamma - 1) (hydro2d/main.cxx:306) and the per-state Dt of :318-333.
    q.E_in = p_in / gmo;
    q.E_out = p_out / gmo;
    auto dt_of = [&](T rho, T E) {
        const T inv = T(1) / rho;
        const T u = T(0) * inv;
        const T tmp = u * u + u * u;
        const T e = E * inv - T(0.5) * tmp;
        const T p = gmo * rho * e;
        const T cs = static_cast<T>(sizeof(T) == 8 ? sqrt(static_cast<double>(gamma * p * inv))
                                                   : sqrtf(static_cast<float>(gamma * p * inv)));
        const T un = static_cast<T>(sizeof(T) == 8 ? sqrt(static_cast<double>(tmp)) : sqrtf(static_cast<float>(tmp)));
        return C_min / (cs + un);
    };
    q.dt_in = dt_of(rho_in, q.E_in);
    q.dt_out = dt_of(rho_out, q.E_out);
    return q;
}

template <typename T>
void ic_launch(T* d_rhoE, T* d_mom, int dim, const size_t* nb, const size_t* base, const IcSetup<T>& q, T rho_in,
               T rho_out, unsigned long long* d_flags, hipStream_t s)
{
    const size_t bz = (dim == 3 && base) ? base[2] : 0, nz = dim == 3 ? nb[2] : 0;
    hipLaunchKernelGGL((ic_kernel<T>), dim3(2048), dim3(256), 0, s, d_rhoE, d_mom, dim, static_cast<long long>(nb[0]),
                       static_cast<long long>(nb[1]), static_cast<long long>(nz),
                       static_cast<long long>(base ? base[0] : 0), static_cast<long long>(base ? base[1] : 0),
                       static_cast<long long>(bz), q.c[0], q.c[1], q.c[2], q.radius, rho_in, rho_out, q.E_in, q.E_out,
                       q.dt_in, q.dt_out, d_flags);
}

template <typename T>
int ic_impl(T* d_rhoE, T* d_mom, int dim, const size_t* nb, const size_t* base, const size_t* NB, T rho_in, T rho_out,
            T p_in, T p_out, int ic_mode, double radius_frac, T C_min, T gamma, T* min_dt, void* stream)
{
    clear_error();
    if (!d_rhoE || !d_mom || !nb || !NB || (dim != 2 && dim != 3) || (ic_mode != 0 && ic_mode != 1)) {
        set_error("hydro_ic: bad arguments");
        return HYDRO_ERR_ARG;
    }
    for (int d = 0; d < dim; ++d)
        if (nb[d] < 1 || NB[d] < 1) {
            set_error("hydro_ic: sizes must be >= 1");
            return HYDRO_ERR_ARG;
        }
    if (!grid_fits(dim, nb) || !grid_fits(dim, NB)) {
        set_error("hydro_ic: grid of more than 2^40 cells");
        return HYDRO_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const IcSetup<T> q = ic_setup<T>(dim, NB, rho_in, rho_out, p_in, p_out, ic_mode, radius_frac, C_min, gamma);

    unsigned long long* d_flags = nullptr;
    HYDRO_HIP_OR(hipMalloc(reinterpret_cast<void**>(&d_flags), sizeof(unsigned long long)), HYDRO_ERR_HIP);
    hipError_t e = hipMemsetAsync(d_flags, 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
        ic_launch<T>(d_rhoE, d_mom, dim, nb, base, q, rho_in, rho_out, d_flags, s);
        e = hipGetLastError();
    }
    unsigned long long flags = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_flags);
    if (e != hipSuccess) {
        set_error("hydro_ic: %s", hipGetErrorString(e));
        return HYDRO_ERR_HIP;
    }
    // min_Dt starts at 1e20 and takes std::min over cells (hydro2d/main.cxx:287, 335).
    T m = static_cast<T>(1e20);
    if (flags & 1u) m = (q.dt_in < m) ? q.dt_in : m;
    if (flags & 2u) m = (q.dt_out < m) ? q.dt_out : m;
    if (min_dt) *min_dt = m;
    return 0;
}

}  // namespace

namespace hydro {

int ic_blast_async_f64(double* d_rhoE, double* d_mom, int dim, const size_t* nb, double rho_in, double rho_out,
                       double p_in, double p_out, int ic_mode, double radius_frac, double C_min, double gamma,
                       unsigned long long* d_flags, double* dt_in, double* dt_out, hipStream_t s)
{
    const IcSetup<double> q =
        ic_setup<double>(dim, nb, rho_in, rho_out, p_in, p_out, ic_mode, radius_frac, C_min, gamma);
    ic_launch<double>(d_rhoE, d_mom, dim, nb, nullptr, q, rho_in, rho_out, d_flags, s);
    *dt_in = q.dt_in;
    *dt_out = q.dt_out;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("hydro IC: %s", hipGetErrorString(e));
        return HYDRO_ERR_HIP;
    }
    return 0;
}

}  // namespace hydro

extern "C" int hydro_ic(double* d_rhoE, double* d_mom, int dim, const size_t* nb, const size_t* base, const size_t* NB,
                        double rho_in, double rho_out, double p_in, double p_out, int ic_mode, double radius_frac,
                        double C_min, double gamma, double* min_dt, void* stream)
{
    return ic_impl<double>(d_rhoE, d_mom, dim, nb, base, NB, rho_in, rho_out, p_in, p_out, ic_mode, radius_frac, C_min,
                           gamma, min_dt, stream);
}

extern "C" int hydro_ic_f32(float* d_rhoE, float* d_mom, int dim, const size_t* nb, const size_t* base,
                            const size_t* NB, float rho_in, float rho_out, float p_in, float p_out, int ic_mode,
                            double radius_frac, float C_min, float gamma, float* min_dt, void* stream)
{
    return ic_impl<float>(d_rhoE, d_mom, dim, nb, base, NB, rho_in, rho_out, p_in, p_out, ic_mode, radius_frac, C_min,
                          gamma, min_dt, stream);
}
