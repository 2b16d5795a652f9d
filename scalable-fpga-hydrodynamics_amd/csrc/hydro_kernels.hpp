// Internal launch interface between the HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "hydro_device.hpp"

namespace hydro {

// Periodic boundaries read from the opposite interior plane / line inside the step
// kernel, so that no ghost fill is needed for them (bits of `wraps`).
constexpr int kWrapX = 1, kWrapY = 2, kWrapZ = 4;

// Reduction state of the plain launchers on stream s of the current device (one per
// stream, so that plain steps on different streams can overlap); NULL with the error
// recorded on failure.
StepReduce* plain_reduce_state(hipStream_t s);

// consts = { C_min, gamma, gamma_minus_one, divgamma, K } in the state precision.  When
// ev_a / ev_b are given they are recorded around the step kernel (not the finalise).
int step2d_resident(const void* rhoE, const void* mom, void* rhoE_next, void* mom_next, size_t nb_x, size_t nb_y,
                    int bc_y, int precision, const double* consts, StepControl* ctl, StepReduce* red, double* traj_t,
                    void* traj_dt, hipStream_t s, hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr,
                    int wraps = 0);
int step3d_resident(const void* rhoE, const void* mom, void* rhoE_next, void* mom_next, size_t nb_x, size_t nb_y,
                    size_t nb_z, int bc_y, int bc_z, int precision, const double* consts, StepControl* ctl,
                    StepReduce* red, double* traj_t, void* traj_dt, hipStream_t s, hipEvent_t ev_a = nullptr,
                    hipEvent_t ev_b = nullptr, int wraps = 0);

int ghost2d_any(void* rhoE, void* mom, size_t nb_x, size_t nb_y, int bc_x, int bc_y, int precision,
                const StepControl* ctl, hipStream_t s);
int ghost3d_any(void* rhoE, void* mom, size_t nb_x, size_t nb_y, size_t nb_z, int bc_x, int bc_y, int bc_z,
                int precision, const StepControl* ctl, hipStream_t s);

// Blast IC (hydro_ic's ic_mode 0/1) written asynchronously on `s`, origin at cell 1;
// the occurring states are or-ed into *d_flags (bit 0 inside, bit 1 outside), which the
// caller clears first, and the two states' Dt are returned on the host.
int ic_blast_async_f64(double* d_rhoE, double* d_mom, int dim, const size_t* nb, double rho_in, double rho_out,
                       double p_in, double p_out, int ic_mode, double radius_frac, double C_min, double gamma,
                       unsigned long long* d_flags, double* dt_in, double* dt_out, hipStream_t s);

}  // namespace hydro

struct hydro_run;

namespace hydro {

// Restart a run created by hydro_run_create for a new blast solve, with no allocation
// and no wait: the blast IC of hydro_batch_add_blast (ic_mode 1, D = 1 / (nb + 1),
// C_min = C * min(D)) written into (rhoE, mom) on s, then the loop control (t = 0,
// it = 0, Dt from the IC) set on the device after it, X periodic and Y/Z reflective.
// FP64 or FP32-flux precision; max_it at most run_capacity(r).  The steps then come
// from hydro_run_advance with (rhoE, mom) as buffer a.
int run_restart_blast(hydro_run* r, int dim, const size_t* nb, int precision, double C, double radius_frac,
                      double rho_in, double rho_out, double p_in, double p_out, double max_t, size_t max_it,
                      void* rhoE, void* mom, hipStream_t s);
size_t run_capacity(const hydro_run* r);

// Dynamic LDS that leaves room for exactly `occ` workgroups of Kernel per CU (160 KiB of
// LDS): the middle of the interval (L/(occ+1), L/occ], less the kernel's static LDS, so
// allocation granularity cannot move the count.  Fewer waves in flight per CU keep fewer
// DRAM rows open at once; the step kernels need no LDS of their own for this.
template <auto Kernel>
size_t occupancy_lds(int occ)
{
    if (occ <= 0) return 0;
    static const size_t static_lds = [] {
        hipFuncAttributes fa;
        return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(Kernel)) == hipSuccess ? fa.sharedSizeBytes
                                                                                              : size_t(0);
    }();
    const size_t total = (kLdsPerCu / (occ + 1) + kLdsPerCu / occ) / 2;
    return total > static_lds ? total - static_lds : 0;
}

// Launch geometry knobs (0 = per-precision default); get_*(true) returns the geometry of
// the last launch after the strip adaptation.
void set_rows_2d(int rows);
void set_occ_2d(int occ);
void set_tile_3d(int rows);
int get_rows_2d(bool last);
int get_occ_2d(bool last);
int get_tile_3d(bool last);

}  // namespace hydro
