// 3D first-order finite-volume Euler step for gfx950 (MI355X).
//
// Reference hot path: kernel_hydro3d_fvm (hydro3d/kernel.cxx:145-407) behind the 3D
// `launcher` (hydro3d/kernel.cxx:419-455), called from hydro3d/main.cxx:791-793.
//
// Work decomposition:
//   * a workgroup covers a (Y x Z) tile of TY rows x wz <= 62 columns (z_window_width:
//     the columns of a row spread evenly over the fewest windows) and marches along X
//     over a strip of rows; it has TY + 2 waves, each holding one Y row j0 - 1 + w
//     (w = 0 and TY+1 are halo rows: they load and convert their row but update
//     nothing; tile_row_of_wave places them on two of the four SIMDs, alternating
//     between co-resident workgroups);
//   * within a wave, lane l holds Z column k = wz*kw + l (lanes 0 and wz+1 are halo
//     columns, lanes above wz+1 idle); the Z face (k, k+1) is computed once by lane k
//     and shuffled to lane k+1;
//   * the Y face (j, j+1) is computed once by the wave of row j from the primitives
//     the wave of row j+1 left in LDS, and left in LDS for the wave of row j+1
//     (two workgroup barriers per X step, single-buffered LDS);
//   * the X face between rows i and i+1 is carried in registers along the march;
//   * TY is 6 for FP64 and FP32 state and 10 for FP32-flux (kTY3);
//   * workgroup b runs on XCD b % 8, and when the strips split evenly over the XCDs,
//     XCD x owns the strips [x*S, (x+1)*S) and walks its tiles in groups of G, about
//     the workgroups resident on the XCD at once, each group through all S strips
//     before the next group: a strip's first rows are then read just after the
//     previous strip's last rows by the same tiles, on the same L2 (512^3: 0.5-1%
//     faster than strip-major order over the whole chip).
// rhoE is read and written as 16-byte (rho, E) pairs and uvw as 24-byte triples; lanes
// are consecutive in Z, so every wave row is one contiguous stream per array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "../../include/hydro_hip.h"
#include "hydro_device.hpp"
#include "hydro_kernels.hpp"
#include "hydro_march3d.hpp"
#include "hydro_runtime.hpp"

namespace hydro {

namespace {

struct Geom3 {
    int nb_x, nb_y, nb_z;
    long long zs, zy;  // nb_z + 2, (nb_y + 2) * (nb_z + 2)
    int rows;          // X rows per strip
    int nzw, nyw;      // Z windows, Y tiles
    int wz;            // owned Z columns per window (z_window_width)
    int wrap_x, wrap_y, wrap_z;  // periodic axes read from the opposite interior plane
    int xcd_strips, group;       // strips per XCD, tile group (XCD-major order; 0: strip-major)
};

// TS: state and update precision; TF, M: precision and math policy of the primitives,
// fluxes and signal speeds (see the 2D kernel).
template <typename TS, typename TF>
constexpr bool kFlux3 = std::is_same_v<TS, double> && std::is_same_v<TF, float>;
template <typename TS, typename TF>
constexpr int kTY3 = kFlux3<TS, TF> ? kTYFlux : kTY;  // owned Y rows per workgroup
// Minimum waves per SIMD: 4 for FP64 (two 8-wave workgroups per CU, at most 128 VGPRs),
// 3 for FP32-flux (one 12-wave workgroup), HYDRO_3D_MIN_WAVES_FP32 for FP32 state.
template <typename TS, typename TF>
constexpr int kMinWaves3 = std::is_same_v<TS, float> ? HYDRO_3D_MIN_WAVES_FP32
                                                     : (kFlux3<TS, TF> ? HYDRO_3D_MIN_WAVES_FLUX : HYDRO_3D_MIN_WAVES);

// Rows of loads ahead (HYDRO_3D_PREFETCH) for the FP64 state; the FP32 state's loads are
// issued where they are used (512^3: 1.287 vs 1.320 ms, alternated in one process; FP64
// equal either way, FP32-flux 1% faster with the row ahead).
template <typename TS>
constexpr int kPF3S = std::is_same_v<TS, float> ? 0 : kPF3;

template <typename TS, typename TF, typename M, bool WALL_Y, bool WALL_Z, bool RESIDENT>
__global__ void __launch_bounds__((kTY3<TS, TF> + 2) * kWave, (kMinWaves3<TS, TF>))
    fvm3d_kernel(const TS* __restrict__ rhoE, const TS* __restrict__ uvw, TS* __restrict__ rhoE_next,
                 TS* __restrict__ uvw_next, Geom3 g, Consts3<TS> cs, TS DtDx_arg, TS DtDy_arg, TS DtDz_arg,
                 StepReduce* red, StepControl* ctl, TS* Dt_next, double* traj_t, TS* traj_dt)
{
    using V2 = typename Vec2_3<TS>::type;

    TS DtDx = DtDx_arg, DtDy = DtDy_arg, DtDz = DtDz_arg;
    if constexpr (RESIDENT) {
        if (ctl->done) return;
        const TS dt = static_cast<TS>(ctl->dt);
        DtDx = dt / static_cast<TS>(ctl->D[0]);  // hydro3d/main.cxx:759-761
        DtDy = dt / static_cast<TS>(ctl->D[1]);
        DtDz = dt / static_cast<TS>(ctl->D[2]);
    }
    const TF gamma = static_cast<TF>(cs.gamma), gmo = static_cast<TF>(cs.gmo);
    const TF divgamma = static_cast<TF>(cs.divgamma), K = static_cast<TF>(cs.K);
    const V2* __restrict__ R2 = reinterpret_cast<const V2*>(rhoE);
    const Trip<TS>* __restrict__ M3 = reinterpret_cast<const Trip<TS>*>(uvw);
    V2* __restrict__ RN2 = reinterpret_cast<V2*>(rhoE_next);
    Trip<TS>* __restrict__ MN3 = reinterpret_cast<Trip<TS>*>(uvw_next);

    const int lane = threadIdx.x & (kWave - 1);
    constexpr int TY = kTY3<TS, TF>, NW = TY + 2;
    __shared__ int s_simd[NW];
    const int w = tile_row_of_wave(s_simd);  // tile row of this wave (halo rows placed by SIMD)
    const unsigned per_strip = static_cast<unsigned>(g.nzw) * static_cast<unsigned>(g.nyw);
    int strip;
    unsigned rem;
    if (g.xcd_strips > 0) {
        // XCD-major order (file comment): XCD x = b % 8 takes its u-th workgroup as
        // (tile group, strip, tile in group), the last group holding the remainder.
        const unsigned S = static_cast<unsigned>(g.xcd_strips), G = static_cast<unsigned>(g.group);
        const unsigned x = blockIdx.x % 8u, u = blockIdx.x / 8u;
        const unsigned full = per_strip / G, in_full = full * G * S;
        unsigned s_in, tile;
        if (u < in_full) {
            const unsigned gi = u / (G * S), r = u % (G * S);
            s_in = r / G;
            tile = gi * G + r % G;
        } else {
            const unsigned Gl = per_strip - full * G, r = u - in_full;
            s_in = r / Gl;
            tile = full * G + r % Gl;
        }
        strip = static_cast<int>(x * S + s_in);
        rem = tile;
    } else {
        const unsigned lb = xcd_remap3(blockIdx.x, gridDim.x);
        strip = static_cast<int>(lb / per_strip);
        rem = lb % per_strip;
    }
    const int yt = static_cast<int>(rem / static_cast<unsigned>(g.nzw));
    const int zw = static_cast<int>(rem % static_cast<unsigned>(g.nzw));

    const int j = yt * TY + w;  // row of this wave (j0 - 1 + w with j0 = 1 + yt*TY)
    const int k = zw * g.wz + lane;
    const bool halo_wave = (w == 0) || (w == NW - 1);
    const bool owned = !halo_wave && j <= g.nb_y && lane >= 1 && lane <= g.wz && k <= g.nb_z;
    // Reflective walls as in 2D: a ghost row (wave) / ghost column (lane) loads its
    // interior neighbour and presents its mirror, normal velocity negated, to the face
    // (hydro3d/kernel.cxx:248-252, 269-273, 290-294, 311-315).
    int jl, kl;
    unsigned int ymask = 0u, zmask = 0u;
    if constexpr (WALL_Y) {
        jl = j < 1 ? 1 : (j > g.nb_y ? g.nb_y : j);
        ymask = (j == 0 || j == g.nb_y + 1) ? kSignMask<TF> : 0u;
    } else {
        jl = j <= g.nb_y + 1 ? j : g.nb_y + 1;
        if (g.wrap_y) jl = j == 0 ? g.nb_y : (jl == g.nb_y + 1 ? 1 : jl);
    }
    if constexpr (WALL_Z) {
        kl = k < 1 ? 1 : (k > g.nb_z ? g.nb_z : k);
        zmask = (k == 0 || k == g.nb_z + 1) ? kSignMask<TF> : 0u;
    } else {
        kl = k <= g.nb_z + 1 ? k : g.nb_z + 1;
        if (g.wrap_z) kl = k == 0 ? g.nb_z : (kl == g.nb_z + 1 ? 1 : kl);
    }
    const long long col = static_cast<long long>(jl) * g.zs + kl;
    const int i0 = 1 + strip * g.rows;
    const int i1 = min(i0 + g.rows - 1, g.nb_x);

    __shared__ March3Lds<TF, NW> lds;
    const uint64_t key = march3d<TS, TF, M, kPF3S<TS>>(R2, M3, RN2, MN3, g.zy, col, g.nb_x, i0, i1, w, lane, halo_wave, owned, ymask,
                                            zmask, g.wrap_x != 0, gamma, gmo, divgamma, K, DtDx, DtDy, DtDz, lds, 0);
    block_max_to_slot(red, key);
}

// X ghost planes (hydro3d/main.cxx:374-405), periodic copy or reflective mirror.
template <typename T>
__global__ void ghost_x3d_kernel(T* __restrict__ rhoE, T* __restrict__ uvw, long long nb_x, long long zy, int bc_x,
                                 const StepControl* ctl)
{
    using V2 = typename Vec2_3<T>::type;
    if (ctl && ctl->done) return;
    V2* R2 = reinterpret_cast<V2*>(rhoE);
    Trip<T>* M3 = reinterpret_cast<Trip<T>*>(uvw);
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 2 * zy) return;
    const int side = t >= zy;
    const long long m = side ? t - zy : t;
    const long long dst = side ? (nb_x + 1) * zy + m : m;
    long long src;
    if (bc_x == HYDRO_BC_PERIODIC) src = side ? zy + m : nb_x * zy + m;
    else src = side ? nb_x * zy + m : zy + m;
    R2[dst] = R2[src];
    Trip<T> v = M3[src];
    if (bc_x != HYDRO_BC_PERIODIC) v.x = -v.x;
    M3[dst] = v;
}

// Periodic Y ghosts in every plane: row 0 <- row nb_y, row nb_y+1 <- row 1 (all k).
template <typename T>
__global__ void ghost_y3d_kernel(T* __restrict__ rhoE, T* __restrict__ uvw, long long nb_x, long long nb_y,
                                 long long zs, const StepControl* ctl)
{
    using V2 = typename Vec2_3<T>::type;
    if (ctl && ctl->done) return;
    V2* R2 = reinterpret_cast<V2*>(rhoE);
    Trip<T>* M3 = reinterpret_cast<Trip<T>*>(uvw);
    const long long zy = (nb_y + 2) * zs;
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 2 * (nb_x + 2) * zs) return;
    const long long k = t % zs;
    const long long q = t / zs;
    const int side = static_cast<int>(q & 1);
    const long long i = q >> 1;
    const long long dst = i * zy + (side ? nb_y + 1 : 0) * zs + k;
    const long long src = i * zy + (side ? 1 : nb_y) * zs + k;
    R2[dst] = R2[src];
    M3[dst] = M3[src];
}

// Periodic Z ghosts in every (i, j) line: k = 0 <- k = nb_z, k = nb_z+1 <- k = 1.
template <typename T>
__global__ void ghost_z3d_kernel(T* __restrict__ rhoE, T* __restrict__ uvw, long long nb_x, long long nb_y,
                                 long long nb_z, const StepControl* ctl)
{
    using V2 = typename Vec2_3<T>::type;
    if (ctl && ctl->done) return;
    V2* R2 = reinterpret_cast<V2*>(rhoE);
    Trip<T>* M3 = reinterpret_cast<Trip<T>*>(uvw);
    const long long zs = nb_z + 2;
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long lines = (nb_x + 2) * (nb_y + 2);
    if (t >= 2 * lines) return;
    const long long line = t >> 1;
    const int side = static_cast<int>(t & 1);
    const long long b = line * zs;
    const long long dst = b + (side ? nb_z + 1 : 0);
    const long long src = b + (side ? 1 : nb_z);
    R2[dst] = R2[src];
    M3[dst] = M3[src];
}

// Maximum X rows per strip (no effect on results), set by hydro_tune_set; 0 selects the
// default of 32 rows (512^3 blast: FP64 and FP32 state fastest at 32 of 4..64 rows;
// FP32-flux at its 10-row tiles 2.34 ms at 32 rows, 2.41 ms at 16).  The rows of the
// last launch (after the strip adaptation) are kept for reporting.
std::atomic<int> g_rows_3d{0};
std::atomic<int> g_last_rows_3d{0};
// Workgroups a step should have to fill the chip (4 per CU); smaller grids use shorter
// strips (down to 4 rows) until they have that many.
constexpr long long kFillBlocks3 = 1024;

// Workgroups of one 3D step kernel resident on one XCD at once (the tile group of the
// XCD-major order), from the kernel's occupancy; 0 if it cannot be queried.
template <typename KernelFn>
int resident_per_xcd(KernelFn kernel, int block)
{
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();  // a failed query must not surface as the launch's error
        return 0;
    }
    return per_cu > 0 && cus >= 8 ? per_cu * (cus / 8) : 0;
}

// Argument checks of a 3D step, run by the plain entry points before they bind the
// stream's reduction state and by launch_step3d (see check_step2d).
int check_step3d(const void* rhoE, const void* uvw, const void* rhoE_next, const void* uvw_next, bool resident,
                 const void* Dt_next, size_t nb_x, size_t nb_y, size_t nb_z, int bc_y, int bc_z)
{
    if (!rhoE || !uvw || !rhoE_next || !uvw_next || (!resident && !Dt_next)) {
        set_error("hydro 3D step: null pointer argument");
        return HYDRO_ERR_ARG;
    }
    const size_t nbv[3] = { nb_x, nb_y, nb_z };
    if (nb_x < 1 || nb_y < 1 || nb_z < 1 || nb_x > (1u << 30) || nb_y > (1u << 20) || nb_z > (1u << 20) ||
        !grid_fits(3, nbv)) {
        set_error("hydro 3D step: interior sizes out of range or above 2^40 cells (%zu x %zu x %zu)", nb_x, nb_y,
                  nb_z);
        return HYDRO_ERR_ARG;
    }
    if ((bc_y != HYDRO_BC_PERIODIC && bc_y != HYDRO_BC_REFLECTIVE) ||
        (bc_z != HYDRO_BC_PERIODIC && bc_z != HYDRO_BC_REFLECTIVE)) {
        set_error("hydro 3D step: unknown boundary mode (%d, %d)", bc_y, bc_z);
        return HYDRO_ERR_ARG;
    }
    return 0;
}

template <typename T, typename TF = T, typename M = IeeeMath>
int launch_step3d(const T* rhoE, const T* uvw, T* rhoE_next, T* uvw_next, const Consts3<T>& c, size_t nb_x,
                  size_t nb_y, size_t nb_z, T DtDx, T DtDy, T DtDz, int bc_y, int bc_z, StepReduce* red,
                  StepControl* ctl, T* Dt_next, double* traj_t, T* traj_dt, hipStream_t s, hipEvent_t ev_a = nullptr,
                  hipEvent_t ev_b = nullptr, int wraps = 0)
{
    if (const int rc =
            check_step3d(rhoE, uvw, rhoE_next, uvw_next, ctl != nullptr, Dt_next, nb_x, nb_y, nb_z, bc_y, bc_z))
        return rc;
    Geom3 g;
    g.nb_x = static_cast<int>(nb_x);
    g.nb_y = static_cast<int>(nb_y);
    g.nb_z = static_cast<int>(nb_z);
    g.zs = static_cast<long long>(nb_z) + 2;
    g.zy = (static_cast<long long>(nb_y) + 2) * g.zs;
    g.nzw = z_windows(g.nb_z);
    g.wz = z_window_width(g.nb_z);
    constexpr int TY = kTY3<T, TF>;
    g.nyw = static_cast<int>((nb_y + TY - 1) / TY);
    g.wrap_x = (wraps & kWrapX) ? 1 : 0;
    g.wrap_y = (wraps & kWrapY) ? 1 : 0;
    g.wrap_z = (wraps & kWrapZ) ? 1 : 0;
    const int rows_set = g_rows_3d.load(std::memory_order_relaxed);
    g.rows = rows_set > 0 ? rows_set : 32;
    while (g.rows > 4 && (static_cast<long long>(nb_x) + g.rows - 1) / g.rows * g.nzw * g.nyw < kFillBlocks3) g.rows /= 2;
    const long long nstrips = (static_cast<long long>(nb_x) + g.rows - 1) / g.rows;
    const long long nblocks = nstrips * g.nzw * g.nyw;
    g.xcd_strips = 0;
    g.group = 0;
    if (nstrips % 8 == 0) {
        // One query per kernel instantiation (all wall variants share its occupancy).
        static std::atomic<int> resident{ -1 };
        int r = resident.load(std::memory_order_relaxed);
        if (r < 0) {
            r = resident_per_xcd(fvm3d_kernel<T, TF, M, true, true, true>, (TY + 2) * kWave);
            resident.store(r, std::memory_order_relaxed);
        }
        if (r > 0) {
            g.xcd_strips = static_cast<int>(nstrips / 8);
            g.group = static_cast<int>(std::min<long long>(r, static_cast<long long>(g.nzw) * g.nyw));
        }
    }
    if (nblocks > 0x7fffffffLL) {
        set_error("hydro 3D step: grid too large (%lld workgroups)", nblocks);
        return HYDRO_ERR_ARG;
    }
    const dim3 grid(static_cast<unsigned>(nblocks)), block((TY + 2) * kWave);
    const bool wy = bc_y == HYDRO_BC_REFLECTIVE, wz = bc_z == HYDRO_BC_REFLECTIVE;
    g_last_rows_3d.store(g.rows, std::memory_order_relaxed);
    if (ev_a) HYDRO_HIP_OR(hipEventRecord(ev_a, s), HYDRO_ERR_HIP);
#define HYDRO_L3(WY, WZ, RES)                                                                                 \
    hipLaunchKernelGGL((fvm3d_kernel<T, TF, M, WY, WZ, RES>), grid, block, 0, s, rhoE, uvw, rhoE_next, uvw_next, g, c, \
                       DtDx, DtDy, DtDz, red, ctl, Dt_next, traj_t, traj_dt)
    if (ctl) {
        if (wy && wz) HYDRO_L3(true, true, true);
        else if (wy) HYDRO_L3(true, false, true);
        else if (wz) HYDRO_L3(false, true, true);
        else HYDRO_L3(false, false, true);
    } else {
        if (wy && wz) HYDRO_L3(true, true, false);
        else if (wy) HYDRO_L3(true, false, false);
        else if (wz) HYDRO_L3(false, true, false);
        else HYDRO_L3(false, false, false);
    }
#undef HYDRO_L3
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    if (ev_b) HYDRO_HIP_OR(hipEventRecord(ev_b, s), HYDRO_ERR_HIP);
    if (ctl)
        hipLaunchKernelGGL((finalize_step_kernel<T, true>), dim3(1), dim3(kWave), 0, s, red, c.C_min, ctl, Dt_next,
                           traj_t, traj_dt);
    else
        hipLaunchKernelGGL((finalize_step_kernel<T, false>), dim3(1), dim3(kWave), 0, s, red, c.C_min, ctl, Dt_next,
                           traj_t, traj_dt);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

template <typename T>
int launch_ghost3d(T* rhoE, T* uvw, size_t nb_x, size_t nb_y, size_t nb_z, int bc_x, int bc_y, int bc_z,
                   const StepControl* ctl, hipStream_t s)
{
    const size_t nbv[3] = { nb_x, nb_y, nb_z };
    if (!rhoE || !uvw || nb_x < 1 || nb_y < 1 || nb_z < 1 || !grid_fits(3, nbv)) {
        set_error("hydro ghost fill 3D: bad arguments");
        return HYDRO_ERR_ARG;
    }
    for (int b : { bc_x, bc_y, bc_z })
        if (b != HYDRO_BC_PERIODIC && b != HYDRO_BC_REFLECTIVE) {
            set_error("hydro ghost fill 3D: unknown boundary mode %d", b);
            return HYDRO_ERR_ARG;
        }
    const long long zs = static_cast<long long>(nb_z) + 2;
    const long long zy = (static_cast<long long>(nb_y) + 2) * zs;
    hipLaunchKernelGGL((ghost_x3d_kernel<T>), dim3(static_cast<unsigned>((2 * zy + 255) / 256)), dim3(256), 0, s,
                       rhoE, uvw, static_cast<long long>(nb_x), zy, bc_x, ctl);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    if (bc_y == HYDRO_BC_PERIODIC) {
        const long long n = 2 * (static_cast<long long>(nb_x) + 2) * zs;
        hipLaunchKernelGGL((ghost_y3d_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                           rhoE, uvw, static_cast<long long>(nb_x), static_cast<long long>(nb_y), zs, ctl);
        HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    }
    if (bc_z == HYDRO_BC_PERIODIC) {
        const long long n = 2 * (static_cast<long long>(nb_x) + 2) * (static_cast<long long>(nb_y) + 2);
        hipLaunchKernelGGL((ghost_z3d_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                           rhoE, uvw, static_cast<long long>(nb_x), static_cast<long long>(nb_y),
                           static_cast<long long>(nb_z), ctl);
        HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    }
    return 0;
}


struct EventPair3 {
    hipEvent_t a = nullptr, b = nullptr;
};

EventPair3* timing_events3()
{
    static thread_local EventPair3 ev[64];
    const int dev = current_device();
    if (dev < 0 || dev >= 64) return nullptr;
    EventPair3& e = ev[dev];
    if (!e.a) {
        HYDRO_HIP_OR(hipEventCreate(&e.a), nullptr);
        HYDRO_HIP_OR(hipEventCreate(&e.b), nullptr);
    }
    return &e;
}

template <typename T>
double timed_launcher3d(const T* d_rhoE, const T* d_uvw, T* d_rhoE_next, T* d_uvw_next, T* Dt_next, T C_min, T gamma,
                        T gamma_minus_one, T divgamma, T K, size_t NB_X, size_t NB_Y, size_t NB_Z, T DtDx, T DtDy,
                        T DtDz, void* stream)
{
    clear_error();
    if (const int rc = check_step3d(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, false, Dt_next, NB_X, NB_Y, NB_Z,
                                    HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE))
        return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StepReduce* red = plain_reduce_state(s);
    EventPair3* ev = timing_events3();
    if (!red || !ev) return HYDRO_ERR_HIP;
    const Consts3<T> c{ C_min, gamma, gamma_minus_one, divgamma, K };
    HYDRO_HIP_OR(hipEventRecord(ev->a, s), HYDRO_ERR_HIP);
    int rc = launch_step3d<T>(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, c, NB_X, NB_Y, NB_Z, DtDx, DtDy, DtDz,
                              HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE, red, nullptr, Dt_next, nullptr, nullptr, s);
    if (rc) return rc;
    HYDRO_HIP_OR(hipEventRecord(ev->b, s), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipEventSynchronize(ev->b), HYDRO_ERR_HIP);
    float ms = 0.f;
    HYDRO_HIP_OR(hipEventElapsedTime(&ms, ev->a, ev->b), HYDRO_ERR_HIP);
    return static_cast<double>(ms) * 1e3;
}

}  // namespace

int step3d_resident(const void* rhoE, const void* mom, void* rhoE_next, void* mom_next, size_t nb_x, size_t nb_y,
                    size_t nb_z, int bc_y, int bc_z, int precision, const double* consts, StepControl* ctl,
                    StepReduce* red, double* traj_t, void* traj_dt, hipStream_t s, hipEvent_t ev_a, hipEvent_t ev_b,
                    int wraps)
{
    if (precision == HYDRO_FP32_FLUX) {
        const Consts3<double> c{ consts[0], consts[1], consts[2], consts[3], consts[4] };
        return launch_step3d<double, float, FastMath>(
            static_cast<const double*>(rhoE), static_cast<const double*>(mom), static_cast<double*>(rhoE_next),
            static_cast<double*>(mom_next), c, nb_x, nb_y, nb_z, 0.0, 0.0, 0.0, bc_y, bc_z, red, ctl, nullptr, traj_t,
            static_cast<double*>(traj_dt), s, ev_a, ev_b, wraps);
    }
    if (precision == 64) {
        const Consts3<double> c{ consts[0], consts[1], consts[2], consts[3], consts[4] };
        return launch_step3d<double>(static_cast<const double*>(rhoE), static_cast<const double*>(mom),
                                     static_cast<double*>(rhoE_next), static_cast<double*>(mom_next), c, nb_x, nb_y,
                                     nb_z, 0.0, 0.0, 0.0, bc_y, bc_z, red, ctl, nullptr, traj_t,
                                     static_cast<double*>(traj_dt), s, ev_a, ev_b, wraps);
    }
    const Consts3<float> c{ static_cast<float>(consts[0]), static_cast<float>(consts[1]),
                            static_cast<float>(consts[2]), static_cast<float>(consts[3]),
                            static_cast<float>(consts[4]) };
    return launch_step3d<float>(static_cast<const float*>(rhoE), static_cast<const float*>(mom),
                                static_cast<float*>(rhoE_next), static_cast<float*>(mom_next), c, nb_x, nb_y, nb_z,
                                0.f, 0.f, 0.f, bc_y, bc_z, red, ctl, nullptr, traj_t, static_cast<float*>(traj_dt), s, ev_a,
                                ev_b, wraps);
}

int ghost3d_any(void* rhoE, void* mom, size_t nb_x, size_t nb_y, size_t nb_z, int bc_x, int bc_y, int bc_z,
                int precision, const StepControl* ctl, hipStream_t s)
{
    if (precision == 64 || precision == HYDRO_FP32_FLUX)
        return launch_ghost3d<double>(static_cast<double*>(rhoE), static_cast<double*>(mom), nb_x, nb_y, nb_z, bc_x,
                                      bc_y, bc_z, ctl, s);
    return launch_ghost3d<float>(static_cast<float*>(rhoE), static_cast<float*>(mom), nb_x, nb_y, nb_z, bc_x, bc_y,
                                 bc_z, ctl, s);
}

void set_tile_3d(int rows) { g_rows_3d.store(rows, std::memory_order_relaxed); }

int get_tile_3d(bool last) { return (last ? g_last_rows_3d : g_rows_3d).load(std::memory_order_relaxed); }

}  // namespace hydro

using namespace hydro;

extern "C" double launcher3d(const double* d_rhoE, const double* d_uvw, double* d_rhoE_next, double* d_uvw_next,
                             double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma,
                             double K, size_t NB_X, size_t NB_Y, size_t NB_Z, double DtDx, double DtDy, double DtDz,
                             void* stream)
{
    return timed_launcher3d<double>(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, Dt_next, C_min, gamma, gamma_minus_one,
                                    divgamma, K, NB_X, NB_Y, NB_Z, DtDx, DtDy, DtDz, stream);
}

extern "C" double launcher3d_f32(const float* d_rhoE, const float* d_uvw, float* d_rhoE_next, float* d_uvw_next,
                                 float* Dt_next, float C_min, float gamma, float gamma_minus_one, float divgamma,
                                 float K, size_t NB_X, size_t NB_Y, size_t NB_Z, float DtDx, float DtDy, float DtDz,
                                 void* stream)
{
    return timed_launcher3d<float>(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, Dt_next, C_min, gamma, gamma_minus_one,
                                   divgamma, K, NB_X, NB_Y, NB_Z, DtDx, DtDy, DtDz, stream);
}

extern "C" int hydro_step3d(const double* d_rhoE, const double* d_uvw, double* d_rhoE_next, double* d_uvw_next,
                            double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma,
                            double K, size_t NB_X, size_t NB_Y, size_t NB_Z, double DtDx, double DtDy, double DtDz,
                            int bc_y, int bc_z, void* stream)
{
    clear_error();
    if (const int rc = check_step3d(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, false, Dt_next, NB_X, NB_Y, NB_Z, bc_y,
                                    bc_z))
        return rc;
    StepReduce* red = plain_reduce_state(static_cast<hipStream_t>(stream));
    if (!red) return HYDRO_ERR_HIP;
    const Consts3<double> c{ C_min, gamma, gamma_minus_one, divgamma, K };
    return launch_step3d<double>(d_rhoE, d_uvw, d_rhoE_next, d_uvw_next, c, NB_X, NB_Y, NB_Z, DtDx, DtDy, DtDz, bc_y,
                                 bc_z, red, nullptr, Dt_next, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

extern "C" int hydro_ghost_fill3d(double* d_rhoE, double* d_uvw, size_t NB_X, size_t NB_Y, size_t NB_Z, int bc_x,
                                  int bc_y, int bc_z, void* stream)
{
    clear_error();
    return launch_ghost3d<double>(d_rhoE, d_uvw, NB_X, NB_Y, NB_Z, bc_x, bc_y, bc_z, nullptr,
                                  static_cast<hipStream_t>(stream));
}

extern "C" int hydro_ghost_fill3d_f32(float* d_rhoE, float* d_uvw, size_t NB_X, size_t NB_Y, size_t NB_Z, int bc_x,
                                      int bc_y, int bc_z, void* stream)
{
    clear_error();
    return launch_ghost3d<float>(d_rhoE, d_uvw, NB_X, NB_Y, NB_Z, bc_x, bc_y, bc_z, nullptr,
                                 static_cast<hipStream_t>(stream));
}
