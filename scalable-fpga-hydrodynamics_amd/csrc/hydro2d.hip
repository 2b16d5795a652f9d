// 2D first-order finite-volume Euler step for gfx950 (MI355X).
//
// Reference hot path: kernel_hydro_fvm (hydro2d/kernel.cxx:132-319) behind `launcher`
// (hydro2d/kernel.cxx:331-367), called once per step from hydro2d/main.cxx:734-736.
//
// Work decomposition (the FPGA line buffer becomes registers of a marching wavefront):
//   * a wave (64 lanes) owns a window of 62 consecutive Y columns plus one halo column
//     on each side (lane l <-> column j = 62*w + l) and marches along X over a strip
//     of `rows` rows, one row per iteration;
//   * the X face between rows i and i+1 is computed once and carried in registers to
//     the next iteration (it is fx2 of row i and fx1 of row i+1);
//   * the Y face between columns j and j+1 is computed once by lane j and handed to
//     lane j+1 with a cross-lane shuffle (fy2 of column j, fy1 of column j+1);
//   * primitives and the impedance sqrt(gamma*p*rho) are computed once per cell.
// Each wave is independent (no LDS, no barrier in the loop); halo columns cost 2/64
// of the lanes.  Loads and stores are 16-byte (rho,E) and (rho*u, rho*v) pairs, so a
// wave row is one contiguous 1 KiB stream per array.
//
// Dt: every lane folds the signal speed of its updated cells into a running max
// (as an order key); waves reduce by shuffles, and each workgroup does one 64-bit
// atomic max into one of 64 slots (spread by workgroup index).  A one-wave finalize
// kernel, launched after the step kernel on the same stream, folds the slots, computes
// Dt = C_min / max_speed and clears the slots for the next step.
#include <hip/hip_runtime.h>

#include <atomic>
#include <type_traits>

#include "../../include/hydro_hip.h"
#include "hydro_device.hpp"
#include "hydro_kernels.hpp"
#include "hydro_march2d.hpp"
#include "hydro_runtime.hpp"

namespace hydro {

namespace {

struct Geom2 {
    int nb_x, nb_y;
    long long ys;  // nb_y + 2
    int rows;      // rows per strip
    int nwin;      // column windows
    int nwb;       // workgroups across Y (nwin / kWavesPerBlock, rounded up)
    int wrap_x;    // periodic X read from the opposite interior plane (no X ghost fill)
    int wrap_y;    // periodic Y read from the opposite interior column (no Y ghost fill)
};

// TS: state and update precision; TF, M: precision and math policy of the primitives,
// fluxes and signal speeds.  <double, double, IeeeMath> is the reference FP64 scheme,
// <float, float, IeeeMath> its USE_FLOAT build, <double, float, FastMath> the FP32-flux
// form (FP32 fluxes, increments and update accumulated in FP64).
// Rows of loads ahead (HYDRO_2D_PREFETCH) for the FP64 state; the FP32 state's loads are
// issued where they are used (5.06 vs 5.11 ms on 40000x20000, alternated in one process).
template <typename TS>
constexpr int kPF2 = std::is_same_v<TS, float> ? 0 : kPF;

template <typename TS, typename TF>
constexpr int kMinBlocks2 =
    std::is_same_v<TS, double> && std::is_same_v<TF, double> ? HYDRO_2D_MIN_BLOCKS_FP64 : HYDRO_2D_MIN_BLOCKS;

template <typename TS, typename TF, typename M, bool WALL_Y, bool RESIDENT>
__global__ void __launch_bounds__(kBlock, (kMinBlocks2<TS, TF>))
    fvm2d_kernel(const TS* __restrict__ rhoE, const TS* __restrict__ uv, TS* __restrict__ rhoE_next,
                 TS* __restrict__ uv_next, Geom2 g, Consts<TS> cs, TS DtDx_arg, TS DtDy_arg, StepReduce* red,
                 StepControl* ctl, TS* Dt_next, double* traj_t, TS* traj_dt)
{
    using V2 = typename Vec2<TS>::type;
    TS DtDx = DtDx_arg, DtDy = DtDy_arg;
    if constexpr (RESIDENT) {
        if (ctl->done) return;  // uniform: the loop has ended, this step is a no-op
        const TS dt = static_cast<TS>(ctl->dt);
        DtDx = dt / static_cast<TS>(ctl->D[0]);  // hydro2d/main.cxx:703-704
        DtDy = dt / static_cast<TS>(ctl->D[1]);
    }
    const TF gamma = static_cast<TF>(cs.gamma), gmo = static_cast<TF>(cs.gmo);
    const TF divgamma = static_cast<TF>(cs.divgamma), K = static_cast<TF>(cs.K);
    const V2* __restrict__ R2 = reinterpret_cast<const V2*>(rhoE);
    const V2* __restrict__ M2 = reinterpret_cast<const V2*>(uv);
    V2* __restrict__ RN2 = reinterpret_cast<V2*>(rhoE_next);
    V2* __restrict__ MN2 = reinterpret_cast<V2*>(uv_next);

    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int wb = static_cast<int>(lb % static_cast<unsigned>(g.nwb));
    const int strip = static_cast<int>(lb / static_cast<unsigned>(g.nwb));
    const int win = wb * kWavesPerBlock + wv;
    const int j = win * kOwned + lane;
    const bool owned = (win < g.nwin) && lane >= 1 && lane <= kOwned && j <= g.nb_y;
    int jl;
    unsigned int ymask;
    window_lane2d<TF, WALL_Y>(j, g.nb_y, g.wrap_y != 0, jl, ymask);
    const int i0 = 1 + strip * g.rows;
    const int i1 = min(i0 + g.rows - 1, g.nb_x);

    uint64_t key = 0;
    if (win < g.nwin)
        key = march2d<TS, TF, M, false, kPF2<TS>>(R2, M2, RN2, MN2, g.ys, g.nb_x, i0, i1, j, jl, ymask, owned,
                                                  g.wrap_x != 0, gamma, gmo, divgamma, K, DtDx, DtDy, key);

    block_max_to_slot(red, key);
}

// The FP64 step's memory schedule without its arithmetic (a measurement ceiling): the
// same workgroups, windows, strips, block order and occupancy as fvm2d_kernel, the same
// loads (rows i0-1 .. i1+1, kPF rows ahead of the row being stored) and the same stores
// (the owned cells of every row); each owned cell of (rhoE, mom) is copied unchanged to
// (rhoE_next, mom_next).  The halo rows' loads are kept alive by a select on a bit
// pattern no state holds, so they cost what the step's do.
__global__ void __launch_bounds__(kBlock, HYDRO_2D_MIN_BLOCKS_FP64)
    pattern2d_kernel(const double2* __restrict__ R2, const double2* __restrict__ M2, double2* __restrict__ RN2,
                     double2* __restrict__ MN2, Geom2 g)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int wb = static_cast<int>(lb % static_cast<unsigned>(g.nwb));
    const int strip = static_cast<int>(lb / static_cast<unsigned>(g.nwb));
    const int win = wb * kWavesPerBlock + wv;
    if (win >= g.nwin) return;
    const int j = win * kOwned + lane;
    const int jl = j <= g.nb_y + 1 ? j : g.nb_y + 1;
    const bool owned = lane >= 1 && lane <= kOwned && j <= g.nb_y;
    const int i0 = 1 + strip * g.rows;
    const int i1 = min(i0 + g.rows - 1, g.nb_x);
    const long long ys = g.ys;
    auto clamp = [&](int r) { return min(r, i1 + 1); };
    constexpr unsigned long long kNever = 0x7ff4dead0badf00dull;  // a NaN payload no state holds
    double2 qr[kPF + 1], qm[kPF + 1];
    const double2 hr = R2[static_cast<long long>(i0 - 1) * ys + jl];  // the halo row behind the strip
    double2 rc = R2[static_cast<long long>(i0) * ys + jl], mc = M2[static_cast<long long>(i0) * ys + jl];
    const double2 hm = M2[static_cast<long long>(i0 - 1) * ys + jl];
#pragma unroll
    for (int q = 0; q < kPF; ++q) {
        const long long n = static_cast<long long>(clamp(i0 + 1 + q)) * ys + jl;
        qr[q] = R2[n];
        qm[q] = M2[n];
    }
    const bool never = __double_as_longlong(hr.x) == static_cast<long long>(kNever) ||
                       __double_as_longlong(hm.x) == static_cast<long long>(kNever);
    for (int i = i0; i <= i1; ++i) {
        const long long nn = static_cast<long long>(clamp(i + 1 + kPF)) * ys + jl;
        qr[kPF] = R2[nn];
        qm[kPF] = M2[nn];
        if (owned) {
            const long long n = static_cast<long long>(i) * ys + j;
            double2 o, p;  // component-wise selects (a select of whole vectors goes through scratch)
            o.x = never ? hr.x : rc.x;
            o.y = never ? hr.y : rc.y;
            p.x = never ? hm.x : mc.x;
            p.y = never ? hm.y : mc.y;
            RN2[n] = o;
            MN2[n] = p;
        }
        rc = qr[0];
        mc = qm[0];
#pragma unroll
        for (int q = 0; q < kPF; ++q) {
            qr[q] = qr[q + 1];
            qm[q] = qm[q + 1];
        }
    }
}

// Ghost fill: X planes (periodic copy or reflective mirror), hydro2d/main.cxx:349-378.
template <typename T>
__global__ void ghost_x2d_kernel(T* __restrict__ rhoE, T* __restrict__ uv, long long nb_x, long long ys, int bc_x,
                                 const StepControl* ctl)
{
    using V2 = typename Vec2<T>::type;
    if (ctl && ctl->done) return;
    V2* R2 = reinterpret_cast<V2*>(rhoE);
    V2* M2 = reinterpret_cast<V2*>(uv);
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 2 * ys) return;
    const int side = t >= ys;  // 0: plane 0, 1: plane nb_x+1
    const long long j = side ? t - ys : t;
    long long dst, src;
    if (bc_x == HYDRO_BC_PERIODIC) {
        dst = side ? (nb_x + 1) * ys + j : j;
        src = side ? ys + j : nb_x * ys + j;
    } else {
        dst = side ? (nb_x + 1) * ys + j : j;
        src = side ? nb_x * ys + j : ys + j;
    }
    R2[dst] = R2[src];
    V2 m = M2[src];
    if (bc_x != HYDRO_BC_PERIODIC) m.x = -m.x;
    M2[dst] = m;
}

// Periodic Y ghosts in every plane (after the X planes): column 0 <- column nb_y,
// column nb_y+1 <- column 1.
template <typename T>
__global__ void ghost_y2d_kernel(T* __restrict__ rhoE, T* __restrict__ uv, long long nb_x, long long nb_y,
                                 const StepControl* ctl)
{
    using V2 = typename Vec2<T>::type;
    if (ctl && ctl->done) return;
    V2* R2 = reinterpret_cast<V2*>(rhoE);
    V2* M2 = reinterpret_cast<V2*>(uv);
    const long long ys = nb_y + 2;
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 2 * (nb_x + 2)) return;
    const long long i = t >> 1;
    const int side = static_cast<int>(t & 1);
    const long long dst = i * ys + (side ? nb_y + 1 : 0);
    const long long src = i * ys + (side ? 1 : nb_y);
    R2[dst] = R2[src];
    M2[dst] = M2[src];
}

// Launch geometry (no effect on results): maximum rows per strip and workgroups per CU,
// set by hydro_tune_set; 0 selects the per-precision defaults below.  The geometry of
// the last launch (after the strip adaptation) is kept for reporting.
std::atomic<int> g_rows_2d{0}, g_occ_2d{0};
std::atomic<int> g_last_rows_2d{0}, g_last_occ_2d{0};

// Defaults from sweeps of the 40000x20000 blast on MI355X (strips of 8..64 rows x 3..8
// workgroups per CU, the configurations alternated inside one process): short strips and
// fewer waves in flight keep fewer DRAM rows open at once.  With one row of loads ahead
// (HYDRO_2D_PREFETCH), 12-row strips with 3 workgroups per CU (FP64, FP32-flux) and, with
// no row ahead, 16-row strips with 6 (FP32 state) are the fastest: FP64 9.55 ms against
// 9.79 at 4 per CU and 10.15 at 5; FP32 state 5.04 ms against 5.06 at 5 and 5.12 at 7;
// 64-row strips at full occupancy ran 7-19% slower.
template <typename T, typename TF>
constexpr int default_rows_2d()
{
    return std::is_same_v<T, float> ? 16 : 12;
}
template <typename T, typename TF>
constexpr int default_occ_2d()
{
    if constexpr (std::is_same_v<T, float>) return 6;
    return 3;
}

// Workgroups a step should have to fill the chip (8 per CU on 256 CUs); smaller grids
// use shorter strips (down to 8 rows) until they have that many.
constexpr long long kFillBlocks2 = 2048;

// Argument checks of a 2D step, run by the plain entry points before they bind the
// stream's reduction state (so that a rejected call leaves nothing behind) and by
// launch_step2d.  `resident`: Dt lives in the run's control block, not in Dt_next.
int check_step2d(const void* rhoE, const void* uv, const void* rhoE_next, const void* uv_next, bool resident,
                 const void* Dt_next, size_t nb_x, size_t nb_y, int bc_y)
{
    if (!rhoE || !uv || !rhoE_next || !uv_next || (!resident && !Dt_next)) {
        set_error("hydro 2D step: null pointer argument");
        return HYDRO_ERR_ARG;
    }
    const size_t nbv[2] = { nb_x, nb_y };
    if (nb_x < 1 || nb_y < 1 || nb_x > (1u << 30) || nb_y > (1u << 30) || !grid_fits(2, nbv)) {
        set_error("hydro 2D step: interior sizes must be in [1, 2^30] with at most 2^40 cells, got %zu x %zu", nb_x,
                  nb_y);
        return HYDRO_ERR_ARG;
    }
    if (bc_y != HYDRO_BC_PERIODIC && bc_y != HYDRO_BC_REFLECTIVE) {
        set_error("hydro 2D step: unknown bc_y %d", bc_y);
        return HYDRO_ERR_ARG;
    }
    return 0;
}

template <typename T, typename TF = T, typename M = IeeeMath>
int launch_step2d(const T* rhoE, const T* uv, T* rhoE_next, T* uv_next, const Consts<T>& c, size_t nb_x, size_t nb_y,
                  T DtDx, T DtDy, int bc_y, StepReduce* red, StepControl* ctl, T* Dt_next, double* traj_t, T* traj_dt,
                  hipStream_t s, hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr, int wraps = 0)
{
    if (const int rc = check_step2d(rhoE, uv, rhoE_next, uv_next, ctl != nullptr, Dt_next, nb_x, nb_y, bc_y)) return rc;
    Geom2 g;
    g.nb_x = static_cast<int>(nb_x);
    g.nb_y = static_cast<int>(nb_y);
    g.ys = static_cast<long long>(nb_y) + 2;
    g.nwin = static_cast<int>((nb_y + kOwned - 1) / kOwned);
    g.nwb = (g.nwin + kWavesPerBlock - 1) / kWavesPerBlock;
    g.wrap_x = (wraps & kWrapX) ? 1 : 0;
    g.wrap_y = (wraps & kWrapY) ? 1 : 0;
    const int rows_set = g_rows_2d.load(std::memory_order_relaxed);
    const int occ_set = g_occ_2d.load(std::memory_order_relaxed);
    const int occ = occ_set > 0 ? occ_set : default_occ_2d<T, TF>();
    g.rows = rows_set > 0 ? rows_set : default_rows_2d<T, TF>();
    while (g.rows > 8 && (static_cast<long long>(nb_x) + g.rows - 1) / g.rows * g.nwb < kFillBlocks2) g.rows /= 2;
    const long long nstrips = (static_cast<long long>(nb_x) + g.rows - 1) / g.rows;
    const long long nblocks = nstrips * g.nwb;
    if (nblocks > 0x7fffffffLL) {
        set_error("hydro 2D step: grid too large (%lld workgroups)", nblocks);
        return HYDRO_ERR_ARG;
    }
    const dim3 grid(static_cast<unsigned>(nblocks)), block(kBlock);
    const bool wall = bc_y == HYDRO_BC_REFLECTIVE;
    g_last_rows_2d.store(g.rows, std::memory_order_relaxed);
    g_last_occ_2d.store(occ, std::memory_order_relaxed);
    if (ev_a) HYDRO_HIP_OR(hipEventRecord(ev_a, s), HYDRO_ERR_HIP);
    auto go = [&](auto wall_c, auto resident_c) {
        constexpr auto kernel = fvm2d_kernel<T, TF, M, decltype(wall_c)::value, decltype(resident_c)::value>;
        const T dx = decltype(resident_c)::value ? T(0) : DtDx, dy = decltype(resident_c)::value ? T(0) : DtDy;
        hipLaunchKernelGGL(kernel, grid, block, occupancy_lds<kernel>(occ), s, rhoE, uv, rhoE_next, uv_next, g, c,
                           dx, dy, red, ctl, Dt_next, traj_t, traj_dt);
    };
    using yes = std::true_type;
    using no = std::false_type;
    if (ctl) {
        if (wall)
            go(yes{}, yes{});
        else
            go(no{}, yes{});
    } else {
        if (wall)
            go(yes{}, no{});
        else
            go(no{}, no{});
    }
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
int launch_ghost2d(T* rhoE, T* uv, size_t nb_x, size_t nb_y, int bc_x, int bc_y, const StepControl* ctl,
                   hipStream_t s)
{
    const size_t nbv[2] = { nb_x, nb_y };
    if (!rhoE || !uv || nb_x < 1 || nb_y < 1 || !grid_fits(2, nbv)) {
        set_error("hydro ghost fill 2D: bad arguments");
        return HYDRO_ERR_ARG;
    }
    if ((bc_x != HYDRO_BC_PERIODIC && bc_x != HYDRO_BC_REFLECTIVE) ||
        (bc_y != HYDRO_BC_PERIODIC && bc_y != HYDRO_BC_REFLECTIVE)) {
        set_error("hydro ghost fill 2D: unknown boundary mode (%d, %d)", bc_x, bc_y);
        return HYDRO_ERR_ARG;
    }
    const long long ys = static_cast<long long>(nb_y) + 2;
    const unsigned nbx = static_cast<unsigned>((2 * ys + 255) / 256);
    hipLaunchKernelGGL((ghost_x2d_kernel<T>), dim3(nbx), dim3(256), 0, s, rhoE, uv, static_cast<long long>(nb_x), ys,
                       bc_x, ctl);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    if (bc_y == HYDRO_BC_PERIODIC) {
        const unsigned nby = static_cast<unsigned>((2 * (static_cast<long long>(nb_x) + 2) + 255) / 256);
        hipLaunchKernelGGL((ghost_y2d_kernel<T>), dim3(nby), dim3(256), 0, s, rhoE, uv, static_cast<long long>(nb_x),
                           static_cast<long long>(nb_y), ctl);
        HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    }
    return 0;
}


struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

EventPair* timing_events()
{
    static thread_local EventPair ev[64];
    const int dev = current_device();
    if (dev < 0 || dev >= 64) return nullptr;
    EventPair& e = ev[dev];
    if (!e.a) {
        HYDRO_HIP_OR(hipEventCreate(&e.a), nullptr);
        HYDRO_HIP_OR(hipEventCreate(&e.b), nullptr);
    }
    return &e;
}

template <typename T>
double timed_launcher2d(const T* d_rhoE, const T* d_uv, T* d_rhoE_next, T* d_uv_next, T* Dt_next, T C_min, T gamma,
                        T gamma_minus_one, T divgamma, T K, size_t NB_X, size_t NB_Y, T DtDx, T DtDy, void* stream)
{
    clear_error();
    if (const int rc = check_step2d(d_rhoE, d_uv, d_rhoE_next, d_uv_next, false, Dt_next, NB_X, NB_Y,
                                    HYDRO_BC_REFLECTIVE))
        return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StepReduce* red = plain_reduce_state(s);
    EventPair* ev = timing_events();
    if (!red || !ev) return HYDRO_ERR_HIP;
    const Consts<T> c{ C_min, gamma, gamma_minus_one, divgamma, K };
    HYDRO_HIP_OR(hipEventRecord(ev->a, s), HYDRO_ERR_HIP);
    int rc = launch_step2d<T>(d_rhoE, d_uv, d_rhoE_next, d_uv_next, c, NB_X, NB_Y, DtDx, DtDy, HYDRO_BC_REFLECTIVE,
                              red, nullptr, Dt_next, nullptr, nullptr, s);
    if (rc) return rc;
    HYDRO_HIP_OR(hipEventRecord(ev->b, s), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipEventSynchronize(ev->b), HYDRO_ERR_HIP);
    float ms = 0.f;
    HYDRO_HIP_OR(hipEventElapsedTime(&ms, ev->a, ev->b), HYDRO_ERR_HIP);
    return static_cast<double>(ms) * 1e3;
}

}  // namespace

// ---- entry points used by the resident run loop and the batched forms -------------

int step2d_resident(const void* rhoE, const void* mom, void* rhoE_next, void* mom_next, size_t nb_x, size_t nb_y,
                    int bc_y, int precision, const double* consts, StepControl* ctl, StepReduce* red, double* traj_t,
                    void* traj_dt, hipStream_t s, hipEvent_t ev_a, hipEvent_t ev_b, int wraps)
{
    if (precision == HYDRO_FP32_FLUX) {
        const Consts<double> c{ consts[0], consts[1], consts[2], consts[3], consts[4] };
        return launch_step2d<double, float, FastMath>(static_cast<const double*>(rhoE), static_cast<const double*>(mom),
                                                      static_cast<double*>(rhoE_next), static_cast<double*>(mom_next), c,
                                                      nb_x, nb_y, 0.0, 0.0, bc_y, red, ctl, nullptr, traj_t,
                                                      static_cast<double*>(traj_dt), s, ev_a, ev_b, wraps);
    }
    if (precision == 64) {
        const Consts<double> c{ consts[0], consts[1], consts[2], consts[3], consts[4] };
        return launch_step2d<double>(static_cast<const double*>(rhoE), static_cast<const double*>(mom),
                                     static_cast<double*>(rhoE_next), static_cast<double*>(mom_next), c, nb_x, nb_y,
                                     0.0, 0.0, bc_y, red, ctl, nullptr, traj_t, static_cast<double*>(traj_dt), s, ev_a,
                                     ev_b, wraps);
    }
    const Consts<float> c{ static_cast<float>(consts[0]), static_cast<float>(consts[1]),
                           static_cast<float>(consts[2]), static_cast<float>(consts[3]),
                           static_cast<float>(consts[4]) };
    return launch_step2d<float>(static_cast<const float*>(rhoE), static_cast<const float*>(mom),
                                static_cast<float*>(rhoE_next), static_cast<float*>(mom_next), c, nb_x, nb_y, 0.f, 0.f,
                                bc_y, red, ctl, nullptr, traj_t, static_cast<float*>(traj_dt), s, ev_a, ev_b, wraps);
}

int ghost2d_any(void* rhoE, void* mom, size_t nb_x, size_t nb_y, int bc_x, int bc_y, int precision,
                const StepControl* ctl, hipStream_t s)
{
    if (precision == 64 || precision == HYDRO_FP32_FLUX)
        return launch_ghost2d<double>(static_cast<double*>(rhoE), static_cast<double*>(mom), nb_x, nb_y, bc_x, bc_y,
                                      ctl, s);
    return launch_ghost2d<float>(static_cast<float*>(rhoE), static_cast<float*>(mom), nb_x, nb_y, bc_x, bc_y, ctl, s);
}

void set_rows_2d(int rows) { g_rows_2d.store(rows, std::memory_order_relaxed); }

int get_rows_2d(bool last)
{
    return (last ? g_last_rows_2d : g_rows_2d).load(std::memory_order_relaxed);
}

void set_occ_2d(int occ) { g_occ_2d.store(occ, std::memory_order_relaxed); }

int get_occ_2d(bool last)
{
    return (last ? g_last_occ_2d : g_occ_2d).load(std::memory_order_relaxed);
}

}  // namespace hydro

using namespace hydro;

extern "C" double launcher(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next,
                           double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma,
                           double K, size_t NB_X, size_t NB_Y, double DtDx, double DtDy, void* stream)
{
    return timed_launcher2d<double>(d_rhoE, d_uv, d_rhoE_next, d_uv_next, Dt_next, C_min, gamma, gamma_minus_one,
                                    divgamma, K, NB_X, NB_Y, DtDx, DtDy, stream);
}

extern "C" double launcher_f32(const float* d_rhoE, const float* d_uv, float* d_rhoE_next, float* d_uv_next,
                               float* Dt_next, float C_min, float gamma, float gamma_minus_one, float divgamma,
                               float K, size_t NB_X, size_t NB_Y, float DtDx, float DtDy, void* stream)
{
    return timed_launcher2d<float>(d_rhoE, d_uv, d_rhoE_next, d_uv_next, Dt_next, C_min, gamma, gamma_minus_one,
                                   divgamma, K, NB_X, NB_Y, DtDx, DtDy, stream);
}

extern "C" int hydro_step2d(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next,
                            double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma,
                            double K, size_t NB_X, size_t NB_Y, double DtDx, double DtDy, int bc_y, void* stream)
{
    clear_error();
    if (const int rc = check_step2d(d_rhoE, d_uv, d_rhoE_next, d_uv_next, false, Dt_next, NB_X, NB_Y, bc_y)) return rc;
    StepReduce* red = plain_reduce_state(static_cast<hipStream_t>(stream));
    if (!red) return HYDRO_ERR_HIP;
    const Consts<double> c{ C_min, gamma, gamma_minus_one, divgamma, K };
    return launch_step2d<double>(d_rhoE, d_uv, d_rhoE_next, d_uv_next, c, NB_X, NB_Y, DtDx, DtDy, bc_y, red, nullptr,
                                 Dt_next, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

extern "C" int hydro_stream_pattern2d(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next,
                                      size_t NB_X, size_t NB_Y, void* stream)
{
    clear_error();
    if (!d_rhoE || !d_uv || !d_rhoE_next || !d_uv_next) {
        set_error("hydro_stream_pattern2d: null pointer argument");
        return HYDRO_ERR_ARG;
    }
    const size_t nbv[2] = { NB_X, NB_Y };
    if (NB_X < 1 || NB_Y < 1 || NB_X > (1u << 30) || NB_Y > (1u << 30) || !grid_fits(2, nbv)) {
        set_error("hydro_stream_pattern2d: interior sizes must be in [1, 2^30] with at most 2^40 cells, got %zu x %zu",
                  NB_X, NB_Y);
        return HYDRO_ERR_ARG;
    }
    // The FP64 step's geometry (launch_step2d): windows, strip rows, workgroups per CU.
    Geom2 g{};
    g.nb_x = static_cast<int>(NB_X);
    g.nb_y = static_cast<int>(NB_Y);
    g.ys = static_cast<long long>(NB_Y) + 2;
    g.nwin = static_cast<int>((NB_Y + kOwned - 1) / kOwned);
    g.nwb = (g.nwin + kWavesPerBlock - 1) / kWavesPerBlock;
    const int rows_set = g_rows_2d.load(std::memory_order_relaxed);
    const int occ_set = g_occ_2d.load(std::memory_order_relaxed);
    const int occ = occ_set > 0 ? occ_set : default_occ_2d<double, double>();
    g.rows = rows_set > 0 ? rows_set : default_rows_2d<double, double>();
    while (g.rows > 8 && (static_cast<long long>(NB_X) + g.rows - 1) / g.rows * g.nwb < kFillBlocks2) g.rows /= 2;
    const long long nblocks = (static_cast<long long>(NB_X) + g.rows - 1) / g.rows * g.nwb;
    if (nblocks > 0x7fffffffLL) {
        set_error("hydro_stream_pattern2d: grid too large (%lld workgroups)", nblocks);
        return HYDRO_ERR_ARG;
    }
    hipLaunchKernelGGL(pattern2d_kernel, dim3(static_cast<unsigned>(nblocks)), dim3(kBlock),
                       occupancy_lds<pattern2d_kernel>(occ), static_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2*>(d_rhoE), reinterpret_cast<const double2*>(d_uv),
                       reinterpret_cast<double2*>(d_rhoE_next), reinterpret_cast<double2*>(d_uv_next), g);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

extern "C" int hydro_ghost_fill2d(double* d_rhoE, double* d_uv, size_t NB_X, size_t NB_Y, int bc_x, int bc_y,
                                  void* stream)
{
    clear_error();
    return launch_ghost2d<double>(d_rhoE, d_uv, NB_X, NB_Y, bc_x, bc_y, nullptr, static_cast<hipStream_t>(stream));
}

extern "C" int hydro_ghost_fill2d_f32(float* d_rhoE, float* d_uv, size_t NB_X, size_t NB_Y, int bc_x, int bc_y,
                                      void* stream)
{
    clear_error();
    return launch_ghost2d<float>(d_rhoE, d_uv, NB_X, NB_Y, bc_x, bc_y, nullptr, static_cast<hipStream_t>(stream));
}
