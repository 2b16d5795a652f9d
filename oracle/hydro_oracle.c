/*
 * hydro_oracle.c -- CPU restatement of the reference's first-order finite-volume
 * compressible Euler update, used ONLY as test infrastructure (the checker).
 *
 * Nothing in the product links, loads or calls this file.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * It restates, in the reference's exact floating-point operation order:
 *   - convert_to_primitives[_get_U]   hydro2d/kernel.cxx:5-53,  hydro3d/kernel.cxx:5-58
 *   - compute_fluxes                  hydro2d/kernel.cxx:74-118, hydro3d/kernel.cxx:79-130
 *   - kernel_hydro_fvm / _3d_fvm      hydro2d/kernel.cxx:132-319, hydro3d/kernel.cxx:145-407
 *   - X ghost fill (single device)    hydro2d/main.cxx:349-378,  hydro3d/main.cxx:374-405
 *   - blast initial condition         hydro2d/main.cxx:280-338,  hydro3d/main.cxx:296-363
 *   - compensated total mass          hydro2d/main.cxx:237-265,  hydro3d/main.cxx:249-281
 *   - step loop / Dt / t advance      hydro2d/main.cxx:686-815,  hydro3d/main.cxx:742-876
 *   - CSV row format                  hydro2d/main.cxx:166-235,  hydro3d/main.cxx:172-247
 *
 * Build with -ffp-contract=off (no FMA contraction) so every + - * / sqrt is a single
 * IEEE-754 round-to-nearest operation, as in the reference.  The file is compiled
 * twice: REAL=double (suffix _f64) and REAL=float (suffix _f32, the USE_FLOAT build).
 *
 * Pinning (see tests/): initial mass and initial Dt of all six published runs, the
 * published per-iteration t / Dt trajectories, closed forms of the flux, the uniform
 * state, mirror symmetry, periodic translation, the 3D(nb_z=1) == 2D identity, and
 * convergence of the 1D reduction to the exact Riemann solution.  The float build has
 * its own pins (float constants, the invariants and Sod in float, agreement with the
 * double build to float precision), and so does the boundary-mode extension
 * (translation along periodic Y/Z, axis-swap equivalence of reflective X and Y).  The
 * published runs' printed change in mass is not reproduced by this strict operation
 * sequence, nor by any rounding or contraction model tried (fp_model_study.c).
 *
 * Boundary-mode extension (not in the reference, which hard-wires X periodic and
 * Y/Z reflective): bc code 0 = periodic, 1 = reflective.  Reflective Y/Z use the
 * reference's in-kernel wall mirror; periodic Y/Z read ghost cells filled by the ghost
 * fill.  Reflective X fills the X ghost planes with the mirror of the adjacent interior
 * plane (normal momentum negated), which gives the same face state as the wall mirror.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(ORACLE_F32)
typedef float real;
#define NAME(n) n##_f32
#define RSQRT sqrtf
#else
typedef double real;
#define NAME(n) n##_f64
#define RSQRT sqrt
#endif

#define R_ZERO ((real)0)
#define R_HALF ((real)0.5)
#define R_ONE ((real)1)

/* std::max(a, b) == (a < b) ? b : a   (hydro2d/kernel.hpp:36) */
static inline real mmax(real a, real b) { return (a < b) ? b : a; }
/* std::min(a, b) == (b < a) ? b : a */
static inline real mmin(real a, real b) { return (b < a) ? b : a; }

enum { BC_PERIODIC = 0, BC_REFLECTIVE = 1 };

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ---------------------------------------------------------------------------------
 * Primitive variables, hydro2d/kernel.cxx:5-53.  q = (rho, u, v, p).
 * ------------------------------------------------------------------------------- */
static inline void prim2(real rho, real mx, real my, real E, real gmo, real q[4])
{
    real inv = R_ONE / rho;
    real ux = mx * inv;
    real uy = my * inv;
    real tmp = ux * ux + uy * uy;
    real e = E * inv - R_HALF * tmp;
    q[0] = rho;
    q[1] = ux;
    q[2] = uy;
    q[3] = gmo * rho * e;
}

/* hydro3d/kernel.cxx:5-58.  q = (rho, u, v, w, p). */
static inline void prim3(real rho, real mx, real my, real mz, real E, real gmo, real q[5])
{
    real inv = R_ONE / rho;
    real ux = mx * inv;
    real uy = my * inv;
    real uz = mz * inv;
    real tmp = ux * ux + uy * uy + uz * uz;
    real e = E * inv - R_HALF * tmp;
    q[0] = rho;
    q[1] = ux;
    q[2] = uy;
    q[3] = uz;
    q[4] = gmo * rho * e;
}

/* ---------------------------------------------------------------------------------
 * Face flux, hydro2d/kernel.cxx:74-118.  L, R = (rho, u_normal, u_tangential, p).
 * F = (mass, normal momentum, tangential momentum, energy) in the argument frame.
 * ------------------------------------------------------------------------------- */
static inline void flux2(const real L[4], const real R[4], real K, real gamma, real divgamma, real F[4])
{
    real aL = RSQRT(gamma * L[3] * L[0]);
    real aR = RSQRT(gamma * R[3] * R[0]);
    real a = K * mmax(aL, aR);
    real diva = R_ONE / a;
    real us = R_HALF * ((L[1] + R[1]) - diva * (R[3] - L[3]));
    real ps = R_HALF * ((L[3] + R[3]) - a * (R[1] - L[1]));
    const real *W = (us >= R_ZERO) ? L : R;
    real Ur = W[0];
    real Un = W[1] * W[0];
    real Ut = W[2] * W[0];
    real UE = W[3] * divgamma + R_HALF * W[0] * (W[1] * W[1] + W[2] * W[2]);
    F[0] = Ur * us + R_ZERO;
    F[1] = Un * us + ps;
    F[2] = Ut * us + R_ZERO;
    F[3] = UE * us + ps * us;
}

/* hydro3d/kernel.cxx:79-130.  L, R = (rho, u_n, u_t1, u_t2, p). */
static inline void flux3(const real L[5], const real R[5], real K, real gamma, real divgamma, real F[5])
{
    real aL = RSQRT(gamma * L[4] * L[0]);
    real aR = RSQRT(gamma * R[4] * R[0]);
    real a = K * mmax(aL, aR);
    real diva = R_ONE / a;
    real us = R_HALF * ((L[1] + R[1]) - diva * (R[4] - L[4]));
    real ps = R_HALF * ((L[4] + R[4]) - a * (R[1] - L[1]));
    const real *W = (us >= R_ZERO) ? L : R;
    real Ur = W[0];
    real Un = W[1] * W[0];
    real Ut1 = W[2] * W[0];
    real Ut2 = W[3] * W[0];
    real eint = W[4] * divgamma;
    real ekin = R_HALF * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    real UE = eint + ekin;
    F[0] = Ur * us + R_ZERO;
    F[1] = Un * us + ps;
    F[2] = Ut1 * us + R_ZERO;
    F[3] = Ut2 * us + R_ZERO;
    F[4] = UE * us + ps * us;
}

/* Exposed for closed-form flux tests. */
void NAME(oracle_flux2)(const real *L, const real *R, real K, real gamma, real divgamma, real *F)
{
    flux2(L, R, K, gamma, divgamma, F);
}
void NAME(oracle_flux3)(const real *L, const real *R, real K, real gamma, real divgamma, real *F)
{
    flux3(L, R, K, gamma, divgamma, F);
}

/* Signal speed of a conserved state, hydro2d/kernel.cxx:281-313 (and IC, main.cxx:318-332). */
static inline real speed2(real rho, real mx, real my, real E, real gamma, real gmo)
{
    real inv = R_ONE / rho;
    real ux = mx * inv;
    real uy = my * inv;
    real tmp = ux * ux + uy * uy;
    real e = E * inv - R_HALF * tmp;
    real p = gmo * rho * e;
    real cs = RSQRT(gamma * p * inv);
    real un = RSQRT(tmp);
    return cs + un;
}

static inline real speed3(real rho, real mx, real my, real mz, real E, real gamma, real gmo)
{
    real inv = R_ONE / rho;
    real ux = mx * inv;
    real uy = my * inv;
    real uz = mz * inv;
    real tmp = ux * ux + uy * uy + uz * uz;
    real e = E * inv - R_HALF * tmp;
    real p = gmo * rho * e;
    real cs = RSQRT(gamma * p * inv);
    real un = RSQRT(tmp);
    return cs + un;
}

/* ---------------------------------------------------------------------------------
 * One 2D step over a (sub)domain: the reference launcher's contract
 * (hydro2d/kernel.hpp:63-68, kernel.cxx:132-319).  Reads rhoE/uv including the
 * X ghost planes filled by the caller, writes only interior cells of *_next and
 * returns *Dt_next = C_min / max_speed(new state).
 * rhoE[2n] = rho, rhoE[2n+1] = E, uv[2n] = rho*u, uv[2n+1] = rho*v, n = i*(nb_y+2)+j.
 * bc_y: 1 = reflective walls at j=1 and j=nb_y (reference), 0 = read the Y ghosts.
 * ------------------------------------------------------------------------------- */
void NAME(oracle_launcher2d)(const real *rhoE, const real *uv, real *rhoE_next, real *uv_next, real *Dt_next,
                             real C_min, real gamma, real gmo, real divgamma, real K, size_t nb_x, size_t nb_y,
                             real DtDx, real DtDy, int bc_y, int nthreads)
{
    const size_t ys = nb_y + 2;
    const int wall_y = (bc_y == BC_REFLECTIVE);
    real max_speed = R_ZERO;
    set_threads(nthreads);
#pragma omp parallel
    {
        real local_max = R_ZERO;
#pragma omp for schedule(static)
        for (long ii = 1; ii <= (long)nb_x; ++ii) {
            const size_t i = (size_t)ii;
            for (size_t j = 1; j <= nb_y; ++j) {
                const size_t c = i * ys + j;
                const real Ur = rhoE[2 * c], UE = rhoE[2 * c + 1], Uu = uv[2 * c], Uv = uv[2 * c + 1];
                real qc[4], qn[4], L[4], R[4];
                real fx1[4], fx2[4], fy1[4], fy2[4];
                prim2(Ur, Uu, Uv, UE, gmo, qc);

                /* X-: flux(prim(i-1), prim(i)), args (rho, u, v, p)   kernel.cxx:186-196 */
                {
                    const size_t n = c - ys;
                    prim2(rhoE[2 * n], uv[2 * n], uv[2 * n + 1], rhoE[2 * n + 1], gmo, qn);
                    flux2(qn, qc, K, gamma, divgamma, fx1);
                }
                /* X+: flux(prim(i), prim(i+1))   kernel.cxx:199-210 */
                {
                    const size_t n = c + ys;
                    prim2(rhoE[2 * n], uv[2 * n], uv[2 * n + 1], rhoE[2 * n + 1], gmo, qn);
                    flux2(qc, qn, K, gamma, divgamma, fx2);
                }
                /* Y-: args (rho, v, u, p); mirror at j==1   kernel.cxx:213-230 */
                {
                    if (wall_y && j == 1) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = -qc[2]; qn[3] = qc[3];
                    } else {
                        const size_t n = c - 1;
                        prim2(rhoE[2 * n], uv[2 * n], uv[2 * n + 1], rhoE[2 * n + 1], gmo, qn);
                    }
                    L[0] = qn[0]; L[1] = qn[2]; L[2] = qn[1]; L[3] = qn[3];
                    R[0] = qc[0]; R[1] = qc[2]; R[2] = qc[1]; R[3] = qc[3];
                    flux2(L, R, K, gamma, divgamma, fy1);
                }
                /* Y+: mirror at j==nb_y   kernel.cxx:233-251 */
                {
                    if (wall_y && j == nb_y) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = -qc[2]; qn[3] = qc[3];
                    } else {
                        const size_t n = c + 1;
                        prim2(rhoE[2 * n], uv[2 * n], uv[2 * n + 1], rhoE[2 * n + 1], gmo, qn);
                    }
                    L[0] = qc[0]; L[1] = qc[2]; L[2] = qc[1]; L[3] = qc[3];
                    R[0] = qn[0]; R[1] = qn[2]; R[2] = qn[1]; R[3] = qn[3];
                    flux2(L, R, K, gamma, divgamma, fy2);
                }
                /* Update, kernel.cxx:259-278.  Y fluxes are in the (rho, v, u, E) frame. */
                const real x_r = DtDx * (fx1[0] - fx2[0]);
                const real x_u = DtDx * (fx1[1] - fx2[1]);
                const real x_v = DtDx * (fx1[2] - fx2[2]);
                const real x_E = DtDx * (fx1[3] - fx2[3]);
                const real y_r = DtDy * (fy1[0] - fy2[0]);
                const real y_u = DtDy * (fy1[2] - fy2[2]);
                const real y_v = DtDy * (fy1[1] - fy2[1]);
                const real y_E = DtDy * (fy1[3] - fy2[3]);
                const real nr = Ur + x_r + y_r;
                const real nu = Uu + x_u + y_u;
                const real nv = Uv + x_v + y_v;
                const real nE = UE + x_E + y_E;
                rhoE_next[2 * c] = nr;
                uv_next[2 * c] = nu;
                uv_next[2 * c + 1] = nv;
                rhoE_next[2 * c + 1] = nE;
                local_max = mmax(speed2(nr, nu, nv, nE, gamma, gmo), local_max);
            }
        }
#pragma omp critical
        max_speed = mmax(local_max, max_speed);
    }
    *Dt_next = C_min / max_speed;
}

/* ---------------------------------------------------------------------------------
 * One 3D step, hydro3d/kernel.cxx:145-407.  n = i*zy + j*(nb_z+2) + k.
 * rhoE[2n] = rho, rhoE[2n+1] = E; uvw[3n+{0,1,2}] = rho*u, rho*v, rho*w.
 * X faces args (rho,u,v,w,p); Y faces (rho,v,u,w,p); Z faces (rho,w,v,u,p).
 * ------------------------------------------------------------------------------- */
static inline void load_prim3(const real *rhoE, const real *uvw, size_t n, real gmo, real q[5])
{
    prim3(rhoE[2 * n], uvw[3 * n], uvw[3 * n + 1], uvw[3 * n + 2], rhoE[2 * n + 1], gmo, q);
}

void NAME(oracle_launcher3d)(const real *rhoE, const real *uvw, real *rhoE_next, real *uvw_next, real *Dt_next,
                             real C_min, real gamma, real gmo, real divgamma, real K, size_t nb_x, size_t nb_y,
                             size_t nb_z, real DtDx, real DtDy, real DtDz, int bc_y, int bc_z, int nthreads)
{
    const size_t zs = nb_z + 2;
    const size_t zy = (nb_y + 2) * zs;
    const int wall_y = (bc_y == BC_REFLECTIVE);
    const int wall_z = (bc_z == BC_REFLECTIVE);
    real max_speed = R_ZERO;
    set_threads(nthreads);
#pragma omp parallel
    {
        real local_max = R_ZERO;
#pragma omp for schedule(static)
        for (long ii = 1; ii <= (long)nb_x; ++ii) {
            const size_t i = (size_t)ii;
            for (size_t j = 1; j <= nb_y; ++j) {
                for (size_t k = 1; k <= nb_z; ++k) {
                    const size_t c = i * zy + j * zs + k;
                    const real Ur = rhoE[2 * c], UE = rhoE[2 * c + 1];
                    const real Uu = uvw[3 * c], Uv = uvw[3 * c + 1], Uw = uvw[3 * c + 2];
                    real qc[5], qn[5], L[5], R[5];
                    real fx1[5], fx2[5], fy1[5], fy2[5], fz1[5], fz2[5];
                    prim3(Ur, Uu, Uv, Uw, UE, gmo, qc);

                    /* X faces, kernel.cxx:211-236 */
                    load_prim3(rhoE, uvw, c - zy, gmo, qn);
                    flux3(qn, qc, K, gamma, divgamma, fx1);
                    load_prim3(rhoE, uvw, c + zy, gmo, qn);
                    flux3(qc, qn, K, gamma, divgamma, fx2);

                    /* Y-: mirror negates v at j==1, kernel.cxx:239-257 */
                    if (wall_y && j == 1) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = -qc[2]; qn[3] = qc[3]; qn[4] = qc[4];
                    } else {
                        load_prim3(rhoE, uvw, c - zs, gmo, qn);
                    }
                    L[0] = qn[0]; L[1] = qn[2]; L[2] = qn[1]; L[3] = qn[3]; L[4] = qn[4];
                    R[0] = qc[0]; R[1] = qc[2]; R[2] = qc[1]; R[3] = qc[3]; R[4] = qc[4];
                    flux3(L, R, K, gamma, divgamma, fy1);
                    /* Y+: kernel.cxx:260-278 */
                    if (wall_y && j == nb_y) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = -qc[2]; qn[3] = qc[3]; qn[4] = qc[4];
                    } else {
                        load_prim3(rhoE, uvw, c + zs, gmo, qn);
                    }
                    L[0] = qc[0]; L[1] = qc[2]; L[2] = qc[1]; L[3] = qc[3]; L[4] = qc[4];
                    R[0] = qn[0]; R[1] = qn[2]; R[2] = qn[1]; R[3] = qn[3]; R[4] = qn[4];
                    flux3(L, R, K, gamma, divgamma, fy2);

                    /* Z-: mirror negates w at k==1, args (rho, w, v, u, p), kernel.cxx:281-299 */
                    if (wall_z && k == 1) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = qc[2]; qn[3] = -qc[3]; qn[4] = qc[4];
                    } else {
                        load_prim3(rhoE, uvw, c - 1, gmo, qn);
                    }
                    L[0] = qn[0]; L[1] = qn[3]; L[2] = qn[2]; L[3] = qn[1]; L[4] = qn[4];
                    R[0] = qc[0]; R[1] = qc[3]; R[2] = qc[2]; R[3] = qc[1]; R[4] = qc[4];
                    flux3(L, R, K, gamma, divgamma, fz1);
                    /* Z+: kernel.cxx:302-320 */
                    if (wall_z && k == nb_z) {
                        qn[0] = qc[0]; qn[1] = qc[1]; qn[2] = qc[2]; qn[3] = -qc[3]; qn[4] = qc[4];
                    } else {
                        load_prim3(rhoE, uvw, c + 1, gmo, qn);
                    }
                    L[0] = qc[0]; L[1] = qc[3]; L[2] = qc[2]; L[3] = qc[1]; L[4] = qc[4];
                    R[0] = qn[0]; R[1] = qn[3]; R[2] = qn[2]; R[3] = qn[1]; R[4] = qn[4];
                    flux3(L, R, K, gamma, divgamma, fz2);

                    /* Update, kernel.cxx:330-358. */
                    const real x_r = DtDx * (fx1[0] - fx2[0]);
                    const real x_u = DtDx * (fx1[1] - fx2[1]);
                    const real x_v = DtDx * (fx1[2] - fx2[2]);
                    const real x_w = DtDx * (fx1[3] - fx2[3]);
                    const real x_E = DtDx * (fx1[4] - fx2[4]);
                    const real y_r = DtDy * (fy1[0] - fy2[0]);
                    const real y_u = DtDy * (fy1[2] - fy2[2]);
                    const real y_v = DtDy * (fy1[1] - fy2[1]);
                    const real y_w = DtDy * (fy1[3] - fy2[3]);
                    const real y_E = DtDy * (fy1[4] - fy2[4]);
                    const real z_r = DtDz * (fz1[0] - fz2[0]);
                    const real z_u = DtDz * (fz1[3] - fz2[3]);
                    const real z_v = DtDz * (fz1[2] - fz2[2]);
                    const real z_w = DtDz * (fz1[1] - fz2[1]);
                    const real z_E = DtDz * (fz1[4] - fz2[4]);
                    const real nr = Ur + x_r + y_r + z_r;
                    const real nu = Uu + x_u + y_u + z_u;
                    const real nv = Uv + x_v + y_v + z_v;
                    const real nw = Uw + x_w + y_w + z_w;
                    const real nE = UE + x_E + y_E + z_E;
                    rhoE_next[2 * c] = nr;
                    uvw_next[3 * c] = nu;
                    uvw_next[3 * c + 1] = nv;
                    uvw_next[3 * c + 2] = nw;
                    rhoE_next[2 * c + 1] = nE;
                    local_max = mmax(speed3(nr, nu, nv, nw, nE, gamma, gmo), local_max);
                }
            }
        }
#pragma omp critical
        max_speed = mmax(local_max, max_speed);
    }
    *Dt_next = C_min / max_speed;
}

/* ---------------------------------------------------------------------------------
 * Ghost fill.  X periodic is hydro2d/main.cxx:349-378: plane 0 <- plane nb_x and
 * plane nb_x+1 <- plane 1, all j (ghosts included).  X reflective (extension): plane 0
 * <- mirror(plane 1), plane nb_x+1 <- mirror(plane nb_x), rho*u negated.  Y/Z periodic
 * (extension) copy the opposite interior line into the Y/Z ghosts of every plane.
 * ------------------------------------------------------------------------------- */
void NAME(oracle_ghost_fill2d)(real *rhoE, real *uv, size_t nb_x, size_t nb_y, int bc_x, int bc_y)
{
    const size_t ys = nb_y + 2;
    if (bc_x == BC_PERIODIC) {
        memcpy(rhoE, rhoE + 2 * nb_x * ys, 2 * ys * sizeof(real));
        memcpy(uv, uv + 2 * nb_x * ys, 2 * ys * sizeof(real));
        memcpy(rhoE + 2 * (nb_x + 1) * ys, rhoE + 2 * ys, 2 * ys * sizeof(real));
        memcpy(uv + 2 * (nb_x + 1) * ys, uv + 2 * ys, 2 * ys * sizeof(real));
    } else {
        for (size_t j = 0; j < ys; ++j) {
            const size_t pairs[2][2] = { { j, ys + j }, { (nb_x + 1) * ys + j, nb_x * ys + j } };
            for (int s = 0; s < 2; ++s) {
                const size_t d = pairs[s][0], src = pairs[s][1];
                rhoE[2 * d] = rhoE[2 * src];
                rhoE[2 * d + 1] = rhoE[2 * src + 1];
                uv[2 * d] = -uv[2 * src];
                uv[2 * d + 1] = uv[2 * src + 1];
            }
        }
    }
    if (bc_y == BC_PERIODIC) {
        for (size_t i = 0; i < nb_x + 2; ++i) {
            const size_t lo = i * ys, hi = i * ys + nb_y + 1;
            rhoE[2 * lo] = rhoE[2 * (hi - 1)];
            rhoE[2 * lo + 1] = rhoE[2 * (hi - 1) + 1];
            uv[2 * lo] = uv[2 * (hi - 1)];
            uv[2 * lo + 1] = uv[2 * (hi - 1) + 1];
            rhoE[2 * hi] = rhoE[2 * (lo + 1)];
            rhoE[2 * hi + 1] = rhoE[2 * (lo + 1) + 1];
            uv[2 * hi] = uv[2 * (lo + 1)];
            uv[2 * hi + 1] = uv[2 * (lo + 1) + 1];
        }
    }
}

static inline void copy_cell3(real *rhoE, real *uvw, size_t d, size_t s, int neg)
{
    rhoE[2 * d] = rhoE[2 * s];
    rhoE[2 * d + 1] = rhoE[2 * s + 1];
    uvw[3 * d] = neg ? -uvw[3 * s] : uvw[3 * s];
    uvw[3 * d + 1] = uvw[3 * s + 1];
    uvw[3 * d + 2] = uvw[3 * s + 2];
}

void NAME(oracle_ghost_fill3d)(real *rhoE, real *uvw, size_t nb_x, size_t nb_y, size_t nb_z, int bc_x, int bc_y,
                               int bc_z)
{
    const size_t zs = nb_z + 2;
    const size_t zy = (nb_y + 2) * zs;
    if (bc_x == BC_PERIODIC) {
        /* hydro3d/main.cxx:374-405 */
        memcpy(rhoE, rhoE + 2 * nb_x * zy, 2 * zy * sizeof(real));
        memcpy(uvw, uvw + 3 * nb_x * zy, 3 * zy * sizeof(real));
        memcpy(rhoE + 2 * (nb_x + 1) * zy, rhoE + 2 * zy, 2 * zy * sizeof(real));
        memcpy(uvw + 3 * (nb_x + 1) * zy, uvw + 3 * zy, 3 * zy * sizeof(real));
    } else {
        for (size_t m = 0; m < zy; ++m) {
            copy_cell3(rhoE, uvw, m, zy + m, 1);
            copy_cell3(rhoE, uvw, (nb_x + 1) * zy + m, nb_x * zy + m, 1);
        }
    }
    if (bc_y == BC_PERIODIC) {
        for (size_t i = 0; i < nb_x + 2; ++i)
            for (size_t k = 0; k < zs; ++k) {
                copy_cell3(rhoE, uvw, i * zy + k, i * zy + nb_y * zs + k, 0);
                copy_cell3(rhoE, uvw, i * zy + (nb_y + 1) * zs + k, i * zy + zs + k, 0);
            }
    }
    if (bc_z == BC_PERIODIC) {
        for (size_t i = 0; i < nb_x + 2; ++i)
            for (size_t j = 0; j < nb_y + 2; ++j) {
                const size_t b = i * zy + j * zs;
                copy_cell3(rhoE, uvw, b, b + nb_z, 0);
                copy_cell3(rhoE, uvw, b + nb_z + 1, b + 1, 0);
            }
    }
}

/* ---------------------------------------------------------------------------------
 * Blast initial condition, hydro2d/main.cxx:280-338 (ic_mode 0) on a (sub)domain whose
 * first interior cell has global 0-based indices (base_i, base_j).  ic_mode 1 is the
 * parameterised blast (extension): radius = radius_frac * min(NB_axis + 1) cells.
 * Global sizes NB_X, NB_Y define the centre.  Returns the min over cells of
 * C_min / (c_s + |u|) in *min_dt.  Ghost cells are left as found (the caller zeroes).
 * ------------------------------------------------------------------------------- */
static double ic_radius(int ic_mode, double radius_frac, const double *centre, int dim, const size_t *NB)
{
    if (ic_mode == 0) {
        double m = centre[0];
        for (int d = 1; d < dim; ++d)
            if (centre[d] < m) m = centre[d];
        return m / 3;
    }
    size_t n = NB[0];
    for (int d = 1; d < dim; ++d)
        if (NB[d] < n) n = NB[d];
    return radius_frac * (double)(n + 1);
}

void NAME(oracle_ic2d)(real *rhoE, real *uv, size_t nb_x, size_t nb_y, size_t base_i, size_t base_j, size_t NB_X,
                       size_t NB_Y, real rho_in, real rho_out, real p_in, real p_out, real C_min, real gamma,
                       int ic_mode, double radius_frac, real *min_dt)
{
    const real gmo = gamma - R_ONE;
    const double centre[2] = { (double)(NB_X + 1) / 2.0, (double)(NB_Y + 1) / 2.0 };
    const size_t NB[2] = { NB_X, NB_Y };
    const double radius = ic_radius(ic_mode, radius_frac, centre, 2, NB);
    const size_t ys = nb_y + 2;
    real mdt = (real)1e20;
    for (size_t ci = 0; ci < nb_x; ++ci) {
        for (size_t cj = 0; cj < nb_y; ++cj) {
            const size_t i = base_i + ci + 1, j = base_j + cj + 1;
            const double dx = (double)i - centre[0], dy = (double)j - centre[1];
            const double dist = sqrt(dx * dx + dy * dy);
            const size_t n = (ci + 1) * ys + (cj + 1);
            const real Ur = dist <= radius ? rho_in : rho_out;
            const real UE = (dist <= radius ? p_in : p_out) / gmo;
            const real Uu = R_ZERO, Uv = R_ZERO;
            rhoE[2 * n] = Ur;
            rhoE[2 * n + 1] = UE;
            uv[2 * n] = Uu;
            uv[2 * n + 1] = Uv;
            const real ldt = C_min / speed2(Ur, Uu, Uv, UE, gamma, gmo);
            mdt = mmin(mdt, ldt);
        }
    }
    *min_dt = mdt;
}

void NAME(oracle_ic3d)(real *rhoE, real *uvw, size_t nb_x, size_t nb_y, size_t nb_z, size_t base_i, size_t base_j,
                       size_t base_k, size_t NB_X, size_t NB_Y, size_t NB_Z, real rho_in, real rho_out, real p_in,
                       real p_out, real C_min, real gamma, int ic_mode, double radius_frac, real *min_dt)
{
    const real gmo = gamma - R_ONE;
    const double centre[3] = { (double)(NB_X + 1) / 2.0, (double)(NB_Y + 1) / 2.0, (double)(NB_Z + 1) / 2.0 };
    const size_t NB[3] = { NB_X, NB_Y, NB_Z };
    const double radius = ic_radius(ic_mode, radius_frac, centre, 3, NB);
    const size_t zs = nb_z + 2, zy = (nb_y + 2) * zs;
    real mdt = (real)1e20;
    for (size_t ci = 0; ci < nb_x; ++ci) {
        for (size_t cj = 0; cj < nb_y; ++cj) {
            for (size_t ck = 0; ck < nb_z; ++ck) {
                const size_t i = base_i + ci + 1, j = base_j + cj + 1, k = base_k + ck + 1;
                const double dx = (double)i - centre[0], dy = (double)j - centre[1], dz = (double)k - centre[2];
                const double dist = sqrt(dx * dx + dy * dy + dz * dz);
                const size_t n = (ci + 1) * zy + (cj + 1) * zs + (ck + 1);
                const real Ur = dist <= radius ? rho_in : rho_out;
                const real UE = (dist <= radius ? p_in : p_out) / gmo;
                rhoE[2 * n] = Ur;
                rhoE[2 * n + 1] = UE;
                uvw[3 * n] = R_ZERO;
                uvw[3 * n + 1] = R_ZERO;
                uvw[3 * n + 2] = R_ZERO;
                const real ldt = C_min / speed3(Ur, R_ZERO, R_ZERO, R_ZERO, UE, gamma, gmo);
                mdt = mmin(mdt, ldt);
            }
        }
    }
    *min_dt = mdt;
}

/* Mass and initial Dt of the blast IC on an NB_X x NB_Y (x NB_Z) grid without storing
 * it (used to pin the published initial masses at 4e8-8e8 cells).  Follows the IC
 * classification and the compensated sum order of compute_total_mass, per subdomain
 * of an X split into n_sub slabs (2-rank runs) whose doubles are then summed.
 * nb_z == 0 selects 2D. */
void NAME(oracle_ic_mass_dt)(size_t NB_X, size_t NB_Y, size_t NB_Z, size_t n_sub, real C, double *mass, real *dt0)
{
    const int dim = NB_Z ? 3 : 2;
    const real gamma = (real)1.4, gmo = gamma - R_ONE;
    const real DX = R_ONE / (real)(NB_X + 1), DY = R_ONE / (real)(NB_Y + 1);
    real ms = mmin(DX, DY);
    if (dim == 3) ms = mmin(ms, R_ONE / (real)(NB_Z + 1));
    const real C_min = C * ms;
    const double centre[3] = { (double)(NB_X + 1) / 2.0, (double)(NB_Y + 1) / 2.0,
                               (double)(NB_Z + 1) / 2.0 };
    const size_t NB[3] = { NB_X, NB_Y, NB_Z };
    const double radius = ic_radius(0, 0.0, centre, dim, NB);
    const real rho_in = (real)1.0, rho_out = (real)1.2, p_in = (real)10.0, p_out = (real)0.1;
    const real dt_in = C_min / speed2(rho_in, R_ZERO, R_ZERO, p_in / gmo, gamma, gmo);
    const real dt_out = C_min / speed2(rho_out, R_ZERO, R_ZERO, p_out / gmo, gamma, gmo);
    const size_t sub = NB_X / n_sub;
    double total = 0.0;
    int seen_in = 0, seen_out = 0;
    for (size_t s = 0; s < n_sub; ++s) {
        const size_t base = s * sub;
        const size_t nx = (s == n_sub - 1) ? sub + NB_X % n_sub : sub;
        long double tm = 0.0L, cc = 0.0L;
        const size_t nz = dim == 3 ? NB_Z : 1;
        for (size_t ci = 0; ci < nx; ++ci)
            for (size_t j = 1; j <= NB_Y; ++j)
                for (size_t k = 1; k <= nz; ++k) {
                    const size_t i = base + ci + 1;
                    const double dx = (double)i - centre[0], dy = (double)j - centre[1];
                    double d2 = dx * dx + dy * dy;
                    if (dim == 3) {
                        const double dz = (double)k - centre[2];
                        d2 = d2 + dz * dz;
                    }
                    const int in = sqrt(d2) <= radius;
                    seen_in |= in;
                    seen_out |= !in;
                    const long double var = (long double)(in ? rho_in : rho_out);
                    const long double t = tm + var;
                    if (fabsl(tm) >= var) cc += (tm - t) + var;
                    else cc += (var - t) + tm;
                    tm = t;
                }
        total += (double)(tm + cc);
    }
    *mass = total;
    /* min over cells of C_min/speed, in the reference's std::min order (1e20 start) */
    real m = (real)1e20;
    if (seen_in) m = mmin(m, dt_in);
    if (seen_out) m = mmin(m, dt_out);
    (void)dim;
    *dt0 = m;
}

/* compute_total_mass, hydro2d/main.cxx:237-265: Neumaier-style compensated long double
 * sum of interior rho in i-major order. */
double NAME(oracle_mass2d)(const real *rhoE, size_t nb_x, size_t nb_y)
{
    const size_t ys = nb_y + 2;
    long double tm = 0.0L, cc = 0.0L;
    for (size_t i = 1; i <= nb_x; ++i)
        for (size_t j = 1; j <= nb_y; ++j) {
            const long double var = (long double)rhoE[2 * (i * ys + j)];
            const long double t = tm + var;
            if (fabsl(tm) >= var) cc += (tm - t) + var;
            else cc += (var - t) + tm;
            tm = t;
        }
    return (double)(tm + cc);
}

double NAME(oracle_mass3d)(const real *rhoE, size_t nb_x, size_t nb_y, size_t nb_z)
{
    const size_t zs = nb_z + 2, zy = (nb_y + 2) * zs;
    long double tm = 0.0L, cc = 0.0L;
    for (size_t i = 1; i <= nb_x; ++i)
        for (size_t j = 1; j <= nb_y; ++j)
            for (size_t k = 1; k <= nb_z; ++k) {
                const long double var = (long double)rhoE[2 * (i * zy + j * zs + k)];
                const long double t = tm + var;
                if (fabsl(tm) >= var) cc += (tm - t) + var;
                else cc += (var - t) + tm;
                tm = t;
            }
    return (double)(tm + cc);
}

/* The reference's subdomains and ranks (hydro2d/main.cxx:620-657, hydro3d/main.cxx:667-713):
 * an sdx x sdy (x sdz) split, subdomain (a, b, c) with index s = (a*sdy + b)*sdz + c, of
 * size NB/sd on each axis (the last one along an axis adding NB % sd) at base
 * (NB/sd)*index; rank r owns the subdomains with s % ranks == r, in s order.  The mass
 * (compute_total_mass, hydro2d/main.cxx:237-265, hydro3d/main.cxx:249-281) is then, per
 * rank, one compensated long double sum over its subdomains' interior rho, each in
 * i-major order, rounded to double; the ranks' doubles are added in rank order (what
 * MPI_Allreduce(SUM) gives for one or two ranks). */
static void sd_box(size_t nb, size_t sd, size_t idx, size_t *base, size_t *n)
{
    const size_t size = nb / sd;
    *base = size * idx;
    *n = size + (idx == sd - 1 ? nb % sd : 0);
}

double NAME(oracle_mass_sd)(const real *rhoE, int dim, const size_t *nb, const size_t *sd, size_t ranks)
{
    const size_t sdz = dim == 3 ? sd[2] : 1, nsd = sd[0] * sd[1] * sdz;
    const size_t zs = dim == 3 ? nb[2] + 2 : 1, koff = dim == 3 ? 1 : 0;
    double total = 0.0;
    for (size_t r = 0; r < ranks; ++r) {
        long double tm = 0.0L, cc = 0.0L;
        for (size_t s = r; s < nsd; s += ranks) {
            size_t b[3] = { 0, 0, 0 }, n[3] = { 1, 1, 1 };
            sd_box(nb[0], sd[0], s / (sd[1] * sdz), &b[0], &n[0]);
            sd_box(nb[1], sd[1], (s / sdz) % sd[1], &b[1], &n[1]);
            if (dim == 3) sd_box(nb[2], sdz, s % sdz, &b[2], &n[2]);
            for (size_t i = b[0] + 1; i <= b[0] + n[0]; ++i)
                for (size_t j = b[1] + 1; j <= b[1] + n[1]; ++j)
                    for (size_t k = b[2] + koff; k < b[2] + koff + n[2]; ++k) {
                        const long double var = (long double)rhoE[2 * ((i * (nb[1] + 2) + j) * zs + k)];
                        const long double t = tm + var;
                        if (fabsl(tm) >= var) cc += (tm - t) + var;
                        else cc += (var - t) + tm;
                        tm = t;
                    }
        }
        total += (double)(tm + cc);
    }
    return total;
}

/* ---------------------------------------------------------------------------------
 * The step loop of hydro2d/main.cxx:686-815 on one domain: ghost fill, DtDx = Dt/DX,
 * launcher, swap, Dt <- Dt_next, ++it, t += Dt, while (t < max_t && it < max_it).
 * On entry rhoE/uv hold the initial state (ghosts as the caller left them) and
 * *dt holds the initial Dt; on return they hold the final state and the last Dt.
 * traj_t[it] / traj_dt[it] receive the (t, Dt) pair printed by the reference at
 * iteration it (t before the advance, Dt the new one).  Returns the number of steps.
 * scratch_* must be arrays of the same size as rhoE / uv.
 * ------------------------------------------------------------------------------- */
size_t NAME(oracle_run2d)(real *rhoE, real *uv, real *scratch_rhoE, real *scratch_uv, size_t nb_x, size_t nb_y,
                          double max_t, size_t max_it, real C, int bc_x, int bc_y, real *dt, double *t_out,
                          double *traj_t, real *traj_dt, int nthreads)
{
    const real gamma = (real)1.4, K = (real)1.1;
    const real gmo = gamma - R_ONE, divgamma = R_ONE / gmo;
    const real DX = R_ONE / (real)(nb_x + 1), DY = R_ONE / (real)(nb_y + 1);
    const real C_min = C * mmin(DX, DY);
    const size_t n = 2 * (nb_x + 2) * (nb_y + 2);
    real *cur_r = rhoE, *cur_u = uv, *nxt_r = scratch_rhoE, *nxt_u = scratch_uv;
    double t = 0.0;
    size_t it = 0;
    real Dt = *dt;
    while (t < max_t && it < max_it) {
        NAME(oracle_ghost_fill2d)(cur_r, cur_u, nb_x, nb_y, bc_x, bc_y);
        const real DtDx = Dt / DX, DtDy = Dt / DY;
        real dtn;
        NAME(oracle_launcher2d)(cur_r, cur_u, nxt_r, nxt_u, &dtn, C_min, gamma, gmo, divgamma, K, nb_x, nb_y, DtDx,
                                DtDy, bc_y, nthreads);
        real *tr = cur_r, *tu = cur_u;
        cur_r = nxt_r; cur_u = nxt_u; nxt_r = tr; nxt_u = tu;
        Dt = dtn;
        if (traj_t) traj_t[it] = t;
        if (traj_dt) traj_dt[it] = Dt;
        ++it;
        t += (double)Dt;
    }
    if (cur_r != rhoE) {
        memcpy(rhoE, cur_r, n * sizeof(real));
        memcpy(uv, cur_u, n * sizeof(real));
    }
    *dt = Dt;
    *t_out = t;
    return it;
}

size_t NAME(oracle_run3d)(real *rhoE, real *uvw, real *scratch_rhoE, real *scratch_uvw, size_t nb_x, size_t nb_y,
                          size_t nb_z, double max_t, size_t max_it, real C, int bc_x, int bc_y, int bc_z, real *dt,
                          double *t_out, double *traj_t, real *traj_dt, int nthreads)
{
    const real gamma = (real)1.4, K = (real)1.1;
    const real gmo = gamma - R_ONE, divgamma = R_ONE / gmo;
    const real DX = R_ONE / (real)(nb_x + 1), DY = R_ONE / (real)(nb_y + 1), DZ = R_ONE / (real)(nb_z + 1);
    const real C_min = C * mmin(mmin(DX, DY), DZ);
    const size_t cells = (nb_x + 2) * (nb_y + 2) * (nb_z + 2);
    real *cur_r = rhoE, *cur_u = uvw, *nxt_r = scratch_rhoE, *nxt_u = scratch_uvw;
    double t = 0.0;
    size_t it = 0;
    real Dt = *dt;
    while (t < max_t && it < max_it) {
        NAME(oracle_ghost_fill3d)(cur_r, cur_u, nb_x, nb_y, nb_z, bc_x, bc_y, bc_z);
        const real DtDx = Dt / DX, DtDy = Dt / DY, DtDz = Dt / DZ;
        real dtn;
        NAME(oracle_launcher3d)(cur_r, cur_u, nxt_r, nxt_u, &dtn, C_min, gamma, gmo, divgamma, K, nb_x, nb_y, nb_z,
                                DtDx, DtDy, DtDz, bc_y, bc_z, nthreads);
        real *tr = cur_r, *tu = cur_u;
        cur_r = nxt_r; cur_u = nxt_u; nxt_r = tr; nxt_u = tu;
        Dt = dtn;
        if (traj_t) traj_t[it] = t;
        if (traj_dt) traj_dt[it] = Dt;
        ++it;
        t += (double)Dt;
    }
    if (cur_r != rhoE) {
        memcpy(rhoE, cur_r, 2 * cells * sizeof(real));
        memcpy(uvw, cur_u, 3 * cells * sizeof(real));
    }
    *dt = Dt;
    *t_out = t;
    return it;
}

/* ---------------------------------------------------------------------------------
 * CSV, hydro2d/main.cxx:166-235: header "x,y,rho,p,u,v" then, per interior cell in
 * i-major order, "i,j,rho,E,rho*u,rho*v" with std::ostream defaults, which for a
 * double is printf's %g (precision 6) and for size_t is %zu.  Returns 0 on success.
 * ------------------------------------------------------------------------------- */
int NAME(oracle_write_csv2d)(const char *path, const real *rhoE, const real *uv, size_t nb_x, size_t nb_y)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    fputs("x,y,rho,p,u,v\n", f);
    const size_t ys = nb_y + 2;
    for (size_t i = 0; i < nb_x; ++i)
        for (size_t j = 0; j < nb_y; ++j) {
            const size_t n = (i + 1) * ys + (j + 1);
            fprintf(f, "%zu,%zu,%g,%g,%g,%g\n", i, j, (double)rhoE[2 * n], (double)rhoE[2 * n + 1], (double)uv[2 * n],
                    (double)uv[2 * n + 1]);
        }
    return fclose(f);
}

/* The CSV of write_results with subdomains and ranks (hydro2d/main.cxx:166-235,
 * hydro3d/main.cxx:172-247): each rank's text is its subdomains' rows, in s order, each
 * subdomain in i-major order with 0-based global indices, and the ranks' texts follow
 * one another in rank order after the header (MPI_Exscan offsets). */
int NAME(oracle_write_csv_sd)(const char *path, const real *rhoE, const real *mom, int dim, const size_t *nb,
                              const size_t *sd, size_t ranks)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    fputs(dim == 2 ? "x,y,rho,p,u,v\n" : "x,y,z,rho,p,u,v,w\n", f);
    const size_t sdz = dim == 3 ? sd[2] : 1, nsd = sd[0] * sd[1] * sdz;
    const size_t zs = dim == 3 ? nb[2] + 2 : 1;
    for (size_t r = 0; r < ranks; ++r)
        for (size_t s = r; s < nsd; s += ranks) {
            size_t b[3] = { 0, 0, 0 }, n[3] = { 1, 1, 1 };
            sd_box(nb[0], sd[0], s / (sd[1] * sdz), &b[0], &n[0]);
            sd_box(nb[1], sd[1], (s / sdz) % sd[1], &b[1], &n[1]);
            if (dim == 3) sd_box(nb[2], sdz, s % sdz, &b[2], &n[2]);
            for (size_t i = b[0]; i < b[0] + n[0]; ++i)
                for (size_t j = b[1]; j < b[1] + n[1]; ++j)
                    for (size_t k = b[2]; k < b[2] + n[2]; ++k) {
                        if (dim == 2) {
                            const size_t c = (i + 1) * (nb[1] + 2) + (j + 1);
                            fprintf(f, "%zu,%zu,%g,%g,%g,%g\n", i, j, (double)rhoE[2 * c], (double)rhoE[2 * c + 1],
                                    (double)mom[2 * c], (double)mom[2 * c + 1]);
                        } else {
                            const size_t c = ((i + 1) * (nb[1] + 2) + (j + 1)) * zs + (k + 1);
                            fprintf(f, "%zu,%zu,%zu,%g,%g,%g,%g,%g\n", i, j, k, (double)rhoE[2 * c],
                                    (double)rhoE[2 * c + 1], (double)mom[3 * c], (double)mom[3 * c + 1],
                                    (double)mom[3 * c + 2]);
                        }
                    }
        }
    return fclose(f);
}

int NAME(oracle_write_csv3d)(const char *path, const real *rhoE, const real *uvw, size_t nb_x, size_t nb_y,
                             size_t nb_z)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    fputs("x,y,z,rho,p,u,v,w\n", f);
    const size_t zs = nb_z + 2, zy = (nb_y + 2) * zs;
    for (size_t i = 0; i < nb_x; ++i)
        for (size_t j = 0; j < nb_y; ++j)
            for (size_t k = 0; k < nb_z; ++k) {
                const size_t n = (i + 1) * zy + (j + 1) * zs + (k + 1);
                fprintf(f, "%zu,%zu,%zu,%g,%g,%g,%g,%g\n", i, j, k, (double)rhoE[2 * n], (double)rhoE[2 * n + 1],
                        (double)uvw[3 * n], (double)uvw[3 * n + 1], (double)uvw[3 * n + 2]);
            }
    return fclose(f);
}
