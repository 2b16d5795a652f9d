// Command-line driver: the drop-in replacement of hydro2d/main.cxx (built as `hydro`,
// HYDRO_DIM = 2) and hydro3d/main.cxx (built as `hydro3d`, HYDRO_DIM = 3) on one GPU.
//
// Same options (getopt_long_only, hydro2d/main.cxx:451-477, hydro3d/main.cxx:491-521),
// same defaults, same stdout lines and the same CSV files.  The reference's MPI ranks
// become one domain resident in HBM: the X ghost planes are filled on the device
// (hydro_ghost_fill*), Dt is reduced on the device, and the step loop runs on the device
// (hydro_run_*), with the host reading back only at write steps and at the end.
// --sdx does not change the solution (the reference's exact ghost exchange makes an X
// split invisible); with --ranks N it sets, as in a run of the reference on N MPI ranks,
// the order of the CSV rows (each rank's subdomains in turn) and how the printed total
// mass is summed (per rank, then across ranks).  --sdy/--sdz above 1 are refused: the
// reference's kernel makes Y and Z subdomain edges walls, a different problem.
//
// Added options: --ranks N (default 1), --bcx/--bcy/--bcz periodic|reflective (reference: periodic X,
// reflective Y and Z), --precision fp64|fp32 (fp32 = the reference's USE_FLOAT build),
// --sync-steps (one synchronous `launcher` call per step with the host reading Dt, the
// reference's own loop structure), --no-output (no CSV files), --threads N (CSV
// formatting threads).
#include <getopt.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "hydro_hip.h"
#include "hydro_io.hpp"

#ifndef HYDRO_DIM
#define HYDRO_DIM 2
#endif

namespace {

constexpr int kDim = HYDRO_DIM;
constexpr int kMom = kDim;  // momentum components per cell

struct Options {
    size_t sd[3] = { 2, 1, 1 };
    size_t nb[3] = { 200, 200, 100 };
    double max_t = 1;
    size_t max_it = 1000;
    size_t write_interval = 0;
    std::string out = "output";
    int bc[3] = { HYDRO_BC_PERIODIC, HYDRO_BC_REFLECTIVE, HYDRO_BC_REFLECTIVE };
    int precision = 64;
    bool sync_steps = false;
    bool no_output = false;
    int threads = 0;
    size_t ranks = 1;  // MPI ranks of the reference run whose output order and mass sums to follow
};

double now_us()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return static_cast<double>(ts.tv_sec) * 1e6 + static_cast<double>(ts.tv_nsec) / 1e3;
}

void die_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "Error: %s: %s\n", what, hipGetErrorString(e));
        exit(1);
    }
}

void die_hydro(int rc, const char* what)
{
    if (rc != 0) {
        fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, hydro_last_error());
        exit(1);
    }
}

int print_usage(const char* exec, const Options& o)
{
    if (kDim == 2) {
        printf("Hydrodynamics 2D\n");
        printf("Usage : %s [-jkxytiowh]\n", exec);
    } else {
        printf("Hydrodynamics 3D\n");
        printf("Usage : %s [-xyztiowh]\n", exec);
    }
    printf("\n");
    printf("Options : \n");
    printf(" -j Set the number of sub-domains in the X axis. (--sdx) Default : %zu\n", o.sd[0]);
    printf(" -k Set the number of sub-domains in the Y axis. (--sdy) Default : %zu\n", o.sd[1]);
    if (kDim == 3) printf(" -l Set the number of sub-domains in the Z axis. (--sdz) Default : %zu\n", o.sd[2]);
    printf(" -x Set the number of spatial grid points in the X axis. Default : %zu\n", o.nb[0]);
    printf(" -y Set the number of spatial grid points in the Y axis. Default : %zu\n", o.nb[1]);
    if (kDim == 3) printf(" -z Set the number of spatial grid points in the Z axis. Default : %zu\n", o.nb[2]);
    printf(" -t Set the time at which the simulation stops (does not translate to iterations count). Default : "
           "%.2e\n",
           o.max_t);
    printf(" -i Set the max number of iterations. First of iterations or timestep is reached stops the simulation. "
           "Default : %zu\n",
           o.max_it);
    printf(" -o Set the output filename. Default : %s\n", o.out.c_str());
    printf(" -w Write result data each <w> timestep. Default : only write final result\n");
    printf(" --bcx, --bcy%s periodic|reflective  Boundary modes. Default : periodic X, reflective Y%s\n",
           kDim == 3 ? ", --bcz" : "", kDim == 3 ? " and Z" : "");
    printf(" --precision fp64|fp32  State precision (fp32 = the USE_FLOAT build). Default : fp64\n");
    printf(" --sync-steps  One synchronous launcher call per step, Dt read back by the host\n");
    printf(" --no-output  Do not write CSV files\n");
    printf(" --ranks N  Order the CSV rows and sum the printed mass as a run on N MPI ranks. Default : 1\n");
    printf(" -h, --help Show this message and exit\n");
    return 0;
}

int parse_bc(const char* s)
{
    if (!strcmp(s, "periodic")) return HYDRO_BC_PERIODIC;
    if (!strcmp(s, "reflective")) return HYDRO_BC_REFLECTIVE;
    fprintf(stderr, "Unknown boundary mode '%s' (periodic|reflective)\n", s);
    exit(1);
}

// One background file writer: start() runs a job on its own thread; join() waits for it
// and exits the program if the job failed (the reference aborts on an output error too).
class Writer {
public:
    template <typename F>
    void start(F job, const std::string& what)
    {
        join();
        what_ = what;
        failed_ = false;
        th_ = std::thread([this, job] { failed_ = !job(); });
    }
    void join()
    {
        if (!th_.joinable()) return;
        th_.join();
        if (failed_) {
            fprintf(stderr, "Error: cannot write %s\n", what_.c_str());
            exit(1);
        }
    }
    ~Writer()
    {
        if (th_.joinable()) th_.join();
    }

private:
    std::thread th_;
    std::atomic<bool> failed_{ false };
    std::string what_;
};

template <typename T>
struct HostCopy {
    T* re = nullptr;
    T* mo = nullptr;
    Writer writer;  // declared after the buffers: destroyed (joined) before they are freed
    // n_mo = 0: the (rho, E) array only, for the mass sum.
    void alloc(size_t n_re, size_t n_mo)
    {
        if ((!re && hipHostMalloc(reinterpret_cast<void**>(&re), n_re * sizeof(T), hipHostMallocDefault) != hipSuccess) ||
            (n_mo && !mo &&
             hipHostMalloc(reinterpret_cast<void**>(&mo), n_mo * sizeof(T), hipHostMallocDefault) != hipSuccess)) {
            fprintf(stderr, "Error: cannot allocate %.2f GB of pinned host memory\n",
                    static_cast<double>((n_re + n_mo) * sizeof(T)) / 1e9);
            exit(1);
        }
    }
    ~HostCopy()
    {
        writer.join();
        if (re) (void)hipHostFree(re);
        if (mo) (void)hipHostFree(mo);
    }
};

// compute_total_mass, hydro2d/main.cxx:237-265 (hydro3d/main.cxx:249-281): each rank a
// compensated long double sum of the interior rho of its subdomains, in `order` (each box
// i-major), rounded to double; the ranks' doubles added in rank order (MPI_Allreduce's
// sum for one or two ranks).
template <typename T>
double total_mass(const T* rhoE, const size_t* nb, const std::vector<hydro_io::Box>& order)
{
    const size_t zs = kDim == 3 ? nb[2] + 2 : 1;
    const size_t koff = kDim == 3 ? 1 : 0;
    const size_t ys = nb[1] + 2;
    double total = 0.0;
    size_t b = 0;
    while (b < order.size()) {
        const size_t rank = order[b].rank;
        long double tm = 0.0L, c = 0.0L;
        for (; b < order.size() && order[b].rank == rank; ++b) {
            const hydro_io::Box& x = order[b];
            for (size_t i = x.base[0] + 1; i <= x.base[0] + x.n[0]; ++i)
                for (size_t j = x.base[1] + 1; j <= x.base[1] + x.n[1]; ++j)
                    for (size_t k = x.base[2] + koff; k < x.base[2] + koff + x.n[2]; ++k) {
                        const size_t n = (i * ys + j) * zs + k;
                        const long double var = static_cast<long double>(rhoE[2 * n]);
                        const long double t = tm + var;
                        if (fabsl(tm) >= var) c += (tm - t) + var;
                        else c += (var - t) + tm;
                        tm = t;
                    }
        }
        total += static_cast<double>(tm + c);
    }
    return total;
}

struct Timers {
    std::vector<double> h2d, d2h, total, compute, bx, by;
};

struct Stats {
    double mean = 0, sd = 0, mn = 0, mx = 0;
};

// hydro2d/timers.h:24-62: values <= 1e-10 are dropped before the statistics.
Stats print_stats(const std::vector<double>& v, const char* label)
{
    std::vector<double> f;
    for (double x : v)
        if (std::abs(x) > 1e-10) f.push_back(x);
    Stats s;
    if (!f.empty()) {
        double sum = 0;
        s.mn = f[0];
        s.mx = f[0];
        for (double x : f) {
            sum += x;
            s.mn = std::min(s.mn, x);
            s.mx = std::max(s.mx, x);
        }
        s.mean = sum / static_cast<double>(f.size());
        double q = 0;
        for (double x : f) q += (x - s.mean) * (x - s.mean);
        s.sd = sqrt(q / static_cast<double>(f.size()));
    }
    printf("-----------------------------------------------------------\n%s\n", label);
    printf("Average execution time: (mean ± σ)   %.1f us ± %.1f µs\n", s.mean, s.sd);
    printf("                        (min … max)  %.1f us … %.1f µs\n\n", s.mn, s.mx);
    return s;
}

template <typename T>
int run(const Options& o, const char* argv0)
{
    (void)argv0;
    const size_t* nb = o.nb;
    const T one = T(1);
    T D[3] = { one / static_cast<T>(nb[0] + 1), one / static_cast<T>(nb[1] + 1),
               kDim == 3 ? one / static_cast<T>(nb[2] + 1) : T(0) };
    T min_spacing = D[0];
    for (int d = 1; d < kDim; ++d) min_spacing = (D[d] < min_spacing) ? D[d] : min_spacing;  // std::min({..})

    fprintf(stderr, "Configuration : \n  Using Finite Volume method\n");
    fprintf(stderr, "  Spatial Points X axis (nb_x)  : %zu\n  Spatial Points Y axis (nb_y)  : %zu\n", nb[0], nb[1]);
    if (kDim == 3) fprintf(stderr, "  Spatial Points Z axis (nb_z)  : %zu\n", nb[2]);
    fprintf(stderr, "  Max Timestep (max_t)          : %s\n  Max iterations (max_it)       : %zu\n",
            hydro_io::format_general(o.max_t).c_str(), o.max_it);
    fprintf(stderr, "  Step in x (dx)                : %s\n  Step in y (dy)                : %s\n",
            hydro_io::format_general(static_cast<double>(D[0])).c_str(),
            hydro_io::format_general(static_cast<double>(D[1])).c_str());
    if (kDim == 3)
        fprintf(stderr, "  Step in z (dz)                : %s\n",
                hydro_io::format_general(static_cast<double>(D[2])).c_str());
    if (o.write_interval) fprintf(stderr, "  Write interval                : %zu\n", o.write_interval);
    static const char* bcn[2] = { "periodic", "reflective" };
    fprintf(stderr, "  Boundaries                    : X %s, Y %s%s%s\n", bcn[o.bc[0]], bcn[o.bc[1]],
            kDim == 3 ? ", Z " : "", kDim == 3 ? bcn[o.bc[2]] : "");
    fprintf(stderr, "  Precision                     : %s\n", sizeof(T) == 8 ? "fp64" : "fp32");

    int dev = 0;
    die_hip(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    die_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    size_t free_b = 0, total_b = 0;
    die_hip(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    // The footprint in floating point first (hydro2d/main.cxx:546-549): sizes that wrap a
    // size_t product (a negative -x read as size_t, as the reference reads it) then fail
    // the memory check instead of passing it with a wrapped, small product.
    double cells_d = 1.0;
    for (int d = 0; d < kDim; ++d) cells_d *= static_cast<double>(nb[d]) + 2.0;
    const double bytes_d = 2.0 * (2.0 + kMom) * cells_d * static_cast<double>(sizeof(T));
    const double one_gb = static_cast<double>(1ull << 30);
    fprintf(stderr, "\n[rank 0] Device: %s (%s)\n", prop.name, prop.gcnArchName);
    fprintf(stderr, "[rank 0] Compute units              : %d\n", prop.multiProcessorCount);
    fprintf(stderr, "[rank 0] Max Global Memory          : %.2f GB\n", static_cast<double>(total_b) / one_gb);
    fprintf(stderr, "[rank 0] Estimated Required Mem     : %.2f GB\n\n", bytes_d / one_gb);
    if (bytes_d > static_cast<double>(free_b)) {
        fprintf(stderr, "[rank 0] Insufficient device memory. Needed: %.2f GB, Available: %.2f GB\n",
                bytes_d / one_gb, static_cast<double>(free_b) / one_gb);
        return 1;
    }
    size_t cells_g = 1;  // fits: the footprint is below the device's free memory
    for (int d = 0; d < kDim; ++d) cells_g *= nb[d] + 2;
    const size_t n_re = 2 * cells_g, n_mo = static_cast<size_t>(kMom) * cells_g;
    const size_t bytes_needed = 2 * (n_re + n_mo) * sizeof(T);

    // Constants, hydro2d/main.cxx:595-607 (3D: C = 0.4, hydro3d/main.cxx:649).
    const T rho_in = T(1.0), rho_out = T(1.2), p_in = T(10.0), p_out = T(0.1);
    const T C = kDim == 2 ? T(0.5) : T(0.4);
    const T gamma = T(1.4), K = T(1.1);
    const T gmo = gamma - one, divgamma = one / gmo;
    const T C_min = C * min_spacing;

    T* d_re[2] = { nullptr, nullptr };
    T* d_mo[2] = { nullptr, nullptr };
    T* d_dt = nullptr;
    for (int b = 0; b < 2; ++b) {
        die_hip(hipMalloc(reinterpret_cast<void**>(&d_re[b]), n_re * sizeof(T)), "hipMalloc");
        die_hip(hipMalloc(reinterpret_cast<void**>(&d_mo[b]), n_mo * sizeof(T)), "hipMalloc");
        die_hip(hipMemset(d_re[b], 0, n_re * sizeof(T)), "hipMemset");  // host vectors are value-initialised
        die_hip(hipMemset(d_mo[b], 0, n_mo * sizeof(T)), "hipMemset");
    }
    die_hip(hipMalloc(reinterpret_cast<void**>(&d_dt), sizeof(T)), "hipMalloc");
    hipStream_t s = nullptr;

    fprintf(stderr, "Init device arrays ...       ");
    T Dt = 0;
    {
        const size_t NB[3] = { nb[0], nb[1], nb[2] };
        int rc;
        if constexpr (sizeof(T) == 8)
            rc = hydro_ic(d_re[0], d_mo[0], kDim, nb, nullptr, NB, rho_in, rho_out, p_in, p_out, 0, 0.0, C_min, gamma,
                          &Dt, s);
        else
            rc = hydro_ic_f32(d_re[0], d_mo[0], kDim, nb, nullptr, NB, rho_in, rho_out, p_in, p_out, 0, 0.0, C_min,
                              gamma, &Dt, s);
        die_hydro(rc, "hydro_ic");
    }
    fprintf(stderr, "done\n");

    // Host copy of the state for the CSV files and the host mass sum: pinned memory
    // (hipHostMalloc), allocated on first use.  A file is formatted and written by a
    // background thread from this copy while the device goes on stepping; the next
    // fetch (and the end of the run) waits for that writer first.
    HostCopy<T> host;
    auto fetch = [&](int cur) {
        host.writer.join();
        host.alloc(n_re, n_mo);
        die_hip(hipMemcpy(host.re, d_re[cur], n_re * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        die_hip(hipMemcpy(host.mo, d_mo[cur], n_mo * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    };
    // The reference's subdomains in the order its output walks them, for --ranks N.
    const std::vector<hydro_io::Box> order = hydro_io::reference_order(kDim, nb, o.sd, o.ranks);
    // The mass the reference prints, compute_total_mass (hydro2d/main.cxx:237-265): the
    // host compensated long double sums of interior rho, from the (rho, E) array alone
    // when no CSV needs the momenta (--no-output).
    auto host_mass = [&](int cur) {
        if (!o.no_output) {
            fetch(cur);
        } else {
            host.writer.join();
            host.alloc(n_re, 0);
            die_hip(hipMemcpy(host.re, d_re[cur], n_re * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        }
        return total_mass<T>(host.re, nb, order);
    };
    const double cells_i = static_cast<double>(nb[0]) * static_cast<double>(nb[1]) *
                           static_cast<double>(kDim == 3 ? nb[2] : 1);
    auto mass_line = [&](double m0, double m, size_t it, double t) {
        const double change = fabs(m0 - m) / cells_i;
        printf("[it=%4lu] t=%.4e Total density : %.7e, change in mass: %.1e (should be close to 0)\n",
               static_cast<unsigned long>(it), t, m, change);
        return change;
    };
    auto write_csv = [&](size_t it, double t) {
        if (o.no_output) return;
        const std::string fname = o.out + "_" + std::to_string(it) + "_" + hydro_io::format_fixed4(t) + ".csv";
        host.writer.start([fname, &host, &nb, &o, &order] {
            return hydro_io::write_csv<T>(fname, host.re, host.mo, kDim, nb, o.threads, order);
        }, fname);
    };

    // Initial mass: the host compensated sum, as the reference (hydro2d/main.cxx:662).
    const double initial_mass = host_mass(0);
    fprintf(stderr, "\nInit conditions : Diffusion of a 100 times higher pressurized circle at simulation center\n");
    fprintf(stderr, "  Default density      : %.2f inside, %.2f outside\n", static_cast<double>(rho_in),
            static_cast<double>(rho_out));
    fprintf(stderr, "  Default pression     : %.2f inside, %.2f outside\n", static_cast<double>(p_in),
            static_cast<double>(p_out));
    fprintf(stderr, "  Default u_x velocity : 0.00\n  Default u_y velocity : 0.00\n");
    printf("\nInitial total mass       : %.7e\n\n", initial_mass);
    printf("Initial Dt : %.2e\n", static_cast<double>(Dt));
    if (o.write_interval) write_csv(0, 0.0);

    Timers T_;
    const double simu_t1 = now_us();
    double t = 0.0;
    size_t it = 0;
    int cur = 0;

    auto it_line = [&](size_t i, double ti, double dti, double usage_us, double compute_us) {
        printf("[it=%4lu] hydro_fvm compute time t=%.4e (Dt=%.3e) : %.5f s (%.2f ms, %.0f us, %.0f us, %.0f us, %.2f "
               "ms)\n",
               static_cast<unsigned long>(i), ti, dti, usage_us / 1e6, 0.0, 0.0, 0.0, compute_us, 0.0);
    };
    auto want_write = [&](size_t i, double ti) {
        return o.write_interval && i && (i % o.write_interval == 0) && i < o.max_it && ti < o.max_t;
    };

    if (o.sync_steps) {
        // The reference's loop, hydro2d/main.cxx:689-815, one launcher call per step.
        while (t < o.max_t && it < o.max_it) {
            const double u0 = now_us();
            int rc;
            const double g0 = now_us();
            if constexpr (sizeof(T) == 8) {
                rc = kDim == 2 ? hydro_ghost_fill2d(d_re[cur], d_mo[cur], nb[0], nb[1], o.bc[0], o.bc[1], s)
                               : hydro_ghost_fill3d(d_re[cur], d_mo[cur], nb[0], nb[1], nb[2], o.bc[0], o.bc[1],
                                                    o.bc[2], s);
            } else {
                rc = kDim == 2 ? hydro_ghost_fill2d_f32(d_re[cur], d_mo[cur], nb[0], nb[1], o.bc[0], o.bc[1], s)
                               : hydro_ghost_fill3d_f32(d_re[cur], d_mo[cur], nb[0], nb[1], nb[2], o.bc[0], o.bc[1],
                                                        o.bc[2], s);
            }
            die_hydro(rc, "ghost fill");
            die_hip(hipStreamSynchronize(s), "ghost fill");
            T_.bx.push_back(now_us() - g0);
            const T DtDx = Dt / D[0], DtDy = Dt / D[1], DtDz = kDim == 3 ? Dt / D[2] : T(0);
            double us;
            if (o.bc[1] == HYDRO_BC_REFLECTIVE && (kDim == 2 || o.bc[2] == HYDRO_BC_REFLECTIVE)) {
                if constexpr (sizeof(T) == 8)
                    us = kDim == 2 ? launcher(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min, gamma,
                                              gmo, divgamma, K, nb[0], nb[1], DtDx, DtDy, s)
                                   : launcher3d(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min,
                                                gamma, gmo, divgamma, K, nb[0], nb[1], nb[2], DtDx, DtDy, DtDz, s);
                else
                    us = kDim == 2 ? launcher_f32(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min,
                                                  gamma, gmo, divgamma, K, nb[0], nb[1], DtDx, DtDy, s)
                                   : launcher3d_f32(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min,
                                                    gamma, gmo, divgamma, K, nb[0], nb[1], nb[2], DtDx, DtDy, DtDz,
                                                    s);
            } else {
                if constexpr (sizeof(T) != 8) {
                    fprintf(stderr, "Error: --sync-steps with periodic Y/Z needs --precision fp64\n");
                    return 1;
                } else {
                    const double c0 = now_us();
                    rc = kDim == 2 ? hydro_step2d(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min,
                                                  gamma, gmo, divgamma, K, nb[0], nb[1], DtDx, DtDy, o.bc[1], s)
                                   : hydro_step3d(d_re[cur], d_mo[cur], d_re[1 - cur], d_mo[1 - cur], d_dt, C_min,
                                                  gamma, gmo, divgamma, K, nb[0], nb[1], nb[2], DtDx, DtDy, DtDz,
                                                  o.bc[1], o.bc[2], s);
                    die_hydro(rc, "step");
                    die_hip(hipStreamSynchronize(s), "step");
                    us = now_us() - c0;
                }
            }
            if (us < 0) die_hydro(static_cast<int>(us), "launcher");
            cur = 1 - cur;
            T dtn;
            die_hip(hipMemcpy(&dtn, d_dt, sizeof(T), hipMemcpyDeviceToHost), "Dt D2H");
            Dt = dtn;
            if (!std::isfinite(static_cast<double>(Dt))) {
                fprintf(stderr, "Error: Dt became non-finite at iteration %zu (t=%.4e)\n", it, t);
                return 3;
            }
            T_.compute.push_back(us);
            T_.total.push_back(now_us() - u0);
            if (want_write(it, t)) {
                mass_line(initial_mass, host_mass(cur), it, t);
                write_csv(it, t);
            }
            it_line(it, t, static_cast<double>(Dt), T_.total.back(), us);
            ++it;
            t += static_cast<double>(Dt);
        }
    } else {
        double Dd[3] = { static_cast<double>(D[0]), static_cast<double>(D[1]), static_cast<double>(D[2]) };
        hydro_run* r = hydro_run_create(kDim, nb, Dd, static_cast<double>(C_min), static_cast<double>(Dt), o.max_t,
                                        o.max_it, sizeof(T) == 8 ? 64 : 32, o.bc);
        if (!r) die_hydro(-1, "hydro_run_create");
        die_hydro(hydro_run_set_timing(r, 1), "hydro_run_set_timing");
        std::vector<double> tt, td;
        size_t done_it = 0;
        bool done = !(t < o.max_t && it < o.max_it);
        while (!done) {
            // Run up to the next write step (state after step w, 2w, ...), or to the end.
            size_t target = o.max_it;
            if (o.write_interval) {
                const size_t w = o.write_interval;
                const size_t next_w = (done_it / w + 1) * w;  // next iteration index to write
                target = std::min(o.max_it, next_w + 1);
            }
            const double u0 = now_us();
            die_hydro(hydro_run_advance(r, d_re[0], d_mo[0], d_re[1], d_mo[1], target - done_it, s),
                      "hydro_run_advance");
            size_t it_now = 0;
            double t_now = 0, dt_now = 0;
            int dflag = 0, curb = 0;
            const int rc = hydro_run_status(r, &it_now, &t_now, &dt_now, &dflag, &curb, s);
            const double chunk_us = now_us() - u0;
            double kms = 0;
            size_t launches = 0;
            die_hydro(hydro_run_kernel_time(r, &kms, &launches), "hydro_run_kernel_time");
            const size_t n_new = it_now - done_it;
            tt.resize(it_now);
            td.resize(it_now);
            if (n_new) die_hydro(hydro_run_trajectory(r, done_it, n_new, tt.data() + done_it, td.data() + done_it, s),
                                 "hydro_run_trajectory");
            const double per_step_us = n_new ? chunk_us / static_cast<double>(n_new) : 0;
            const double comp_us = launches ? kms * 1e3 / static_cast<double>(launches) : 0;
            for (size_t i = done_it; i < it_now; ++i) {
                T_.total.push_back(per_step_us);
                T_.compute.push_back(comp_us);
                if (i + 1 == it_now && want_write(i, tt[i])) {
                    mass_line(initial_mass, host_mass(curb), i, tt[i]);
                    write_csv(i, tt[i]);
                }
                it_line(i, tt[i], td[i], per_step_us, comp_us);
            }
            if (rc == HYDRO_ERR_NONFINITE) {
                fprintf(stderr, "Error: %s\n", hydro_last_error());
                hydro_run_destroy(r);
                return 3;
            }
            die_hydro(rc, "hydro_run_status");
            done_it = it_now;
            it = it_now;
            t = t_now;
            Dt = static_cast<T>(dt_now);
            cur = curb;
            done = dflag != 0 || n_new == 0;
        }
        hydro_run_destroy(r);
    }
    const double simu_t2 = now_us();

    const double final_mass = host_mass(cur);
    const double final_change = mass_line(initial_mass, final_mass, it, t);
    write_csv(it, t);
    host.writer.join();

    const double t_simu = (simu_t2 - simu_t1) / 1e6;
    printf("\n");
    Stats h2d = print_stats(T_.h2d, "CPU→GPU");
    Stats d2h = print_stats(T_.d2h, "GPU→CPU");
    Stats tot = print_stats(T_.total, "TOTAL usage");
    Stats comp = print_stats(T_.compute, "TOTAL compute");
    Stats bx = print_stats(T_.bx, "boundaries X");
    print_stats(T_.by, "boundaries Y");
    (void)h2d;
    (void)d2h;
    (void)bx;
    printf("Total execution time: %.3f s\n", t_simu);
    printf("\n");
    printf("Subdomains : [1][1]%s = 1 (one resident domain)\n", kDim == 3 ? "[1]" : "");
    printf("Problem size : %.2e\n", cells_i);
    printf("GPU RAM usage : %.2f GB (%.1f %%)\n", static_cast<double>(bytes_needed) / one_gb,
           100.0 * static_cast<double>(bytes_needed) / static_cast<double>(total_b));
    printf("Mass change : %.2e\n", final_change);
    const double total_bytes = 2.0 * (kDim == 2 ? 4 : 5) * cells_i * static_cast<double>(sizeof(T));
    auto gbs = [](double bytes, double us) { return us == 0.0 ? 0.0 : bytes / (us / 1e6) / 1e9; };
    auto mcs = [](double pts, double us) { return us == 0.0 ? 0.0 : pts / (us / 1e6) / 1e6; };
    printf("Estimated Throughput : %.3f GB/s (all code)\n", gbs(total_bytes, tot.mean));
    printf("Estimated Throughput : %.3f GB/s (GPU compute)\n", gbs(total_bytes, comp.mean));
    printf("Performance : %.3f Mc/s (all code)\n", mcs(cells_i, tot.mean));
    printf("Performance : %.3f Mc/s (GPU compute)\n", mcs(cells_i, comp.mean));
    printf("Performance : %.3f GCUPS (GPU compute), %.1f %% of the 8.0 TB/s HBM roofline\n",
           mcs(cells_i, comp.mean) / 1e3, 100.0 * gbs(total_bytes, comp.mean) / 8000.0);

    for (int b = 0; b < 2; ++b) {
        (void)hipFree(d_re[b]);
        (void)hipFree(d_mo[b]);
    }
    (void)hipFree(d_dt);
    return 0;
}

}  // namespace

int main(int argc, char* argv[])
{
    Options o;
    if (kDim == 3) o.nb[1] = 200;
    enum { OPT_BCX = 1000, OPT_BCY, OPT_BCZ, OPT_PREC, OPT_SYNC, OPT_NOOUT, OPT_THREADS, OPT_RANKS };
    const char* short_options = kDim == 2 ? "j:k:x:y:t:i:o:w:h" : "j:k:l:x:y:z:t:i:o:w:h";
    const struct option long_options[] = { { "sdx", required_argument, nullptr, 'j' },
                                           { "sdy", required_argument, nullptr, 'k' },
                                           { "sdz", required_argument, nullptr, 'l' },
                                           { "x_points", required_argument, nullptr, 'x' },
                                           { "y_points", required_argument, nullptr, 'y' },
                                           { "z_points", required_argument, nullptr, 'z' },
                                           { "max_time", required_argument, nullptr, 't' },
                                           { "max_iterations", required_argument, nullptr, 'i' },
                                           { "output_filename", required_argument, nullptr, 'o' },
                                           { "write_interval", required_argument, nullptr, 'w' },
                                           { "bcx", required_argument, nullptr, OPT_BCX },
                                           { "bcy", required_argument, nullptr, OPT_BCY },
                                           { "bcz", required_argument, nullptr, OPT_BCZ },
                                           { "precision", required_argument, nullptr, OPT_PREC },
                                           { "sync-steps", no_argument, nullptr, OPT_SYNC },
                                           { "no-output", no_argument, nullptr, OPT_NOOUT },
                                           { "threads", required_argument, nullptr, OPT_THREADS },
                                           { "ranks", required_argument, nullptr, OPT_RANKS },
                                           { "help", no_argument, nullptr, 'h' },
                                           { nullptr, 0, nullptr, 0 } };
    int opt;
    try {
        while ((opt = getopt_long_only(argc, argv, short_options, long_options, nullptr)) != -1) {
            switch (opt) {
            case 'h': return print_usage(argv[0], o);
            case 'j': o.sd[0] = static_cast<size_t>(std::stoi(optarg)); break;
            case 'k': o.sd[1] = static_cast<size_t>(std::stoi(optarg)); break;
            case 'l': o.sd[2] = static_cast<size_t>(std::stoi(optarg)); break;
            case 'x': o.nb[0] = static_cast<size_t>(std::stoi(optarg)); break;
            case 'y': o.nb[1] = static_cast<size_t>(std::stoi(optarg)); break;
            case 'z': o.nb[2] = static_cast<size_t>(std::stoi(optarg)); break;
            case 't': o.max_t = atof(optarg); break;
            case 'i': o.max_it = static_cast<size_t>(std::stoi(optarg)); break;
            case 'o': o.out = optarg; break;
            case 'w': o.write_interval = static_cast<size_t>(std::stoi(optarg)); break;
            case OPT_BCX: o.bc[0] = parse_bc(optarg); break;
            case OPT_BCY: o.bc[1] = parse_bc(optarg); break;
            case OPT_BCZ: o.bc[2] = parse_bc(optarg); break;
            case OPT_PREC:
                if (!strcmp(optarg, "fp64")) o.precision = 64;
                else if (!strcmp(optarg, "fp32")) o.precision = 32;
                else {
                    fprintf(stderr, "Unknown precision '%s' (fp64|fp32)\n", optarg);
                    return 1;
                }
                break;
            case OPT_SYNC: o.sync_steps = true; break;
            case OPT_NOOUT: o.no_output = true; break;
            case OPT_THREADS: o.threads = std::stoi(optarg); break;
            case OPT_RANKS: {
                const int r = std::stoi(optarg);
                if (r < 1) {
                    fprintf(stderr, "Error: --ranks must be >= 1\n");
                    return 1;
                }
                o.ranks = static_cast<size_t>(r);
                break;
            }
            default: fprintf(stderr, "Unknown option\n"); return 1;
            }
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "Bad option value: %s\n", e.what());
        return 1;
    }
    for (int d = 0; d < kDim; ++d) {
        if (o.nb[d] < 1) {
            fprintf(stderr, "Error: grid sizes must be >= 1\n");
            return 1;
        }
        if (o.sd[d] < 1 || o.sd[d] > (size_t(1) << 30)) {
            fprintf(stderr, "Error: the number of sub-domains along an axis must be in [1, 2^30]\n");
            return 1;
        }
    }
    // The reference's kernel puts its Y (and Z) walls at every subdomain's own edges
    // (hydro2d/kernel.cxx:214, 235: col_pos == 1, col_pos == y_stride - 2), so a split
    // along Y or Z solves separate channels coupled only through Dt; along X the ghost
    // exchange makes a split exact.  One resident domain reproduces X splits only.
    if (o.sd[1] != 1 || (kDim == 3 && o.sd[2] != 1)) {
        fprintf(stderr, "Error: sub-domains along Y%s are not supported: the reference makes their edges walls; "
                        "split along X only (-j)\n", kDim == 3 ? " or Z" : "");
        return 1;
    }
    if (kDim == 2) o.sd[2] = 1;
    return o.precision == 64 ? run<double>(o, argv[0]) : run<float>(o, argv[0]);
}
