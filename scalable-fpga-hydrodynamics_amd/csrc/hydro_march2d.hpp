// The 2D strip march shared by the single-domain step kernel and the batched kernel.
//
// One wave (64 lanes, lane l <-> Y column j) marches along X over rows i0..i1 of its
// column window, reading rows i0-1 .. i1+1:
//   * the X face between rows i and i+1 is computed once and carried in registers to
//     the next row (fx2 of row i is fx1 of row i+1; hydro2d/kernel.cxx:186-210 computes
//     it twice with identical arguments);
//   * the Y face (j, j+1) is computed once by lane j and handed to lane j+1 by a
//     cross-lane shuffle (hydro2d/kernel.cxx:213-251);
//   * primitives and the impedance sqrt(gamma*p*rho) are computed once per cell;
//   * kPF rows of loads are kept in flight ahead of the row being converted.
// The arithmetic of every cell is the reference's (hydro2d/kernel.cxx:132-319) in its
// operation order, so results do not depend on how rows and windows are split.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "hydro_device.hpp"

namespace hydro {

constexpr int kOwned = kWave - 2;  // owned columns per wave window
constexpr int kWavesPerBlock = 4;
// Workgroups per CU the 2D step must fit in registers: 5 (at most 102 VGPRs), and 4 for
// the FP64 scheme (102 VGPRs; at 5 it spills 6 registers, 7% slower).
#ifndef HYDRO_2D_MIN_BLOCKS
#define HYDRO_2D_MIN_BLOCKS 5
#endif
#ifndef HYDRO_2D_MIN_BLOCKS_FP64
#define HYDRO_2D_MIN_BLOCKS_FP64 4
#endif
// Rows in flight ahead of the row being converted (register prefetch ring depth).
#ifndef HYDRO_2D_PREFETCH
#define HYDRO_2D_PREFETCH 3
#endif
constexpr int kPF = HYDRO_2D_PREFETCH;
#ifndef HYDRO_2D_UNROLL
#define HYDRO_2D_UNROLL 0
#endif
// Non-temporal (streaming) stores of the new state.
#ifndef HYDRO_2D_NT_STORE
#define HYDRO_2D_NT_STORE 0
#endif
constexpr int kBlock = kWave * kWavesPerBlock;

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

template <typename T>
__device__ __forceinline__ void store_pair(typename Vec2<T>::type* p, T a, T b)
{
    if constexpr (HYDRO_2D_NT_STORE) {
        typedef T v2t __attribute__((ext_vector_type(2)));
        v2t v;
        v.x = a;
        v.y = b;
        __builtin_nontemporal_store(v, reinterpret_cast<v2t*>(p));
    } else {
        typename Vec2<T>::type v;
        v.x = a;
        v.y = b;
        *p = v;
    }
}


// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2): give each
// XCD a contiguous range of logical blocks so that neighbouring column windows, which
// share halo columns, run on the same L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n)
{
    constexpr unsigned X = 8;
    const unsigned q = n / X, r = n % X;
    const unsigned xcd = b % X, k = b / X;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

template <typename T>
struct Consts {
    T C_min, gamma, gmo, divgamma, K;
};

template <typename T>
__device__ __forceinline__ void load_cell(const typename Vec2<T>::type* __restrict__ re,
                                          const typename Vec2<T>::type* __restrict__ mo, long long n,
                                          typename Vec2<T>::type& a, typename Vec2<T>::type& b)
{
    a = re[n];
    b = mo[n];
}

// Row actually read for row r.  wrap_x: rows 0 and nb_x + 1 are read from rows nb_x and
// 1 (the periodic X ghost planes, hydro2d/main.cxx:349-378, without materialising them);
// otherwise the ghost planes are read as filled by the ghost fill.  r and wrap_x are
// wave-uniform, so this is scalar arithmetic.
__device__ __forceinline__ long long row_of(int r, int nb_x, bool wrap_x)
{
    return wrap_x ? (r == 0 ? nb_x : (r == nb_x + 1 ? 1 : r)) : r;
}

// Per-lane column set-up for a window whose lane 0 is column j = j_first (Y ghost
// columns included).  With reflective Y walls the two ghost columns load their interior
// neighbour and present its mirror (v negated) to the Y face, which is the reference's
// wall state (hydro2d/kernel.cxx:214-229, 234-250): lane j = 0 evaluates
// flux(mirror(c1), c1) for column 1 and lane j = nb_y + 1 hands mirror(c_nb_y) to column
// nb_y.  Negation is exact, so prims(mirror) == mirror(prims).  Periodic Y reads the
// ghost columns filled by the ghost fill, or with wrap_y the opposite interior column.
// Lanes past nb_y + 1 are clamped in bounds and never owned.
template <typename TF, bool WALL_Y>
__device__ __forceinline__ void window_lane2d(int j, int nb_y, bool wrap_y, int& jl, unsigned int& ymask)
{
    ymask = 0u;
    if constexpr (WALL_Y) {
        jl = j < 1 ? 1 : (j > nb_y ? nb_y : j);
        ymask = (j == 0 || j == nb_y + 1) ? kSignMask<TF> : 0u;
    } else {
        jl = j <= nb_y + 1 ? j : nb_y + 1;
        if (wrap_y) jl = j == 0 ? nb_y : (jl == nb_y + 1 ? 1 : jl);
    }
}

// One row of the march: the Y face of row i (lane j, handed to lane j+1), the primitives
// of row i+1 (into Pn) and the X face (i, i+1) (into F2), the update of row i with the
// X face (i-1, i) in F1, its store and the order key of its new signal speed.
template <typename TS, typename TF, typename M>
__device__ __forceinline__ uint64_t row2d(const typename Vec2<TS>::type& rc, const typename Vec2<TS>::type& mc,
                                          const typename Vec2<TS>::type& rn, const typename Vec2<TS>::type& mn,
                                          const Prim2<TF>& Pc, Prim2<TF>& Pn, const TF (&F1)[4], TF (&F2)[4],
                                          typename Vec2<TS>::type* __restrict__ RN2,
                                          typename Vec2<TS>::type* __restrict__ MN2, long long ys, int i, int j,
                                          unsigned int ymask, bool owned, TF gamma, TF gmo, TF divgamma, TF K,
                                          TS DtDx, TS DtDy, uint64_t key)
{
    // Y face (j, j+1) of row i in the (rho, v, u, p) frame (kernel.cxx:213-251),
    // evaluated by lane j and handed to lane j+1.
    Prim2<TF> Sy, Ny;
    Sy.r = Pc.r;
    Sy.un = xor_sign(Pc.ut, ymask);
    Sy.ut = Pc.un;
    Sy.p = Pc.p;
    Sy.a = Pc.a;
    Ny.r = from_next_lane(Sy.r);
    Ny.un = from_next_lane(Sy.un);
    Ny.ut = from_next_lane(Sy.ut);
    Ny.p = from_next_lane(Sy.p);
    Ny.a = from_next_lane(Sy.a);
    TF fyR0, fyR1, fyR2, fyR3;
    face_flux2<TF, M>(Sy, Ny, K, divgamma, fyR0, fyR1, fyR2, fyR3);
    const TF fyL0 = from_prev_lane(fyR0);
    const TF fyL1 = from_prev_lane(fyR1);
    const TF fyL2 = from_prev_lane(fyR2);
    const TF fyL3 = from_prev_lane(fyR3);

    // Row i+1 primitives and the X face (i, i+1).
    Pn.r = static_cast<TF>(rn.x);
    cons_to_prim2<TF, M>(Pn.r, static_cast<TF>(mn.x), static_cast<TF>(mn.y), static_cast<TF>(rn.y), gmo, gamma, Pn.un,
                         Pn.ut, Pn.p, Pn.a);
    face_flux2<TF, M>(Pc, Pn, K, divgamma, F2[0], F2[1], F2[2], F2[3]);

    // Update, kernel.cxx:259-278: U' = (U + x) + y; Y fluxes are in the (rho, v, u, E)
    // frame, so rho*u takes component 2 and rho*v component 1.  Increments and update
    // are in the state precision TS.
    const TS x_r = DtDx * (static_cast<TS>(F1[0]) - static_cast<TS>(F2[0]));
    const TS x_u = DtDx * (static_cast<TS>(F1[1]) - static_cast<TS>(F2[1]));
    const TS x_v = DtDx * (static_cast<TS>(F1[2]) - static_cast<TS>(F2[2]));
    const TS x_E = DtDx * (static_cast<TS>(F1[3]) - static_cast<TS>(F2[3]));
    const TS y_r = DtDy * (static_cast<TS>(fyL0) - static_cast<TS>(fyR0));
    const TS y_u = DtDy * (static_cast<TS>(fyL2) - static_cast<TS>(fyR2));
    const TS y_v = DtDy * (static_cast<TS>(fyL1) - static_cast<TS>(fyR1));
    const TS y_E = DtDy * (static_cast<TS>(fyL3) - static_cast<TS>(fyR3));
    const TS nr = rc.x + x_r + y_r;
    const TS nu = mc.x + x_u + y_u;
    const TS nv = mc.y + x_v + y_v;
    const TS nE = rc.y + x_E + y_E;
    const TF sp = signal_speed2<TF, M>(static_cast<TF>(nr), static_cast<TF>(nu), static_cast<TF>(nv),
                                       static_cast<TF>(nE), gmo, gamma);
    const uint64_t k_new = speed_key(static_cast<double>(sp));
    key = owned ? umax64(key, k_new) : key;
    if (owned) {
        const long long n = static_cast<long long>(i) * ys + j;
        store_pair<TS>(RN2 + n, nr, nE);
        store_pair<TS>(MN2 + n, nu, nv);
    }
    return key;
}

template <typename F, int... P>
__device__ __forceinline__ bool run_phases(F& step, std::integer_sequence<int, P...>)
{
    return (step(std::integral_constant<int, P>{}) && ...);
}

// March rows i0..i1 (i0 <= i1) of one window.  Lane columns: j (global, for stores and
// ownership) and jl (the column loaded); wrap_x as in row_of.  Returns `key` folded with the order keys of
// the new signal speeds of the owned cells.
//
// kPF rows of loads are in flight while a row is computed.  HYDRO_2D_UNROLL: rows
// i .. i+1+kPF live in a ring of kPF+2 register slots and the primitives and X-face
// flux of rows i and i+1 in two slots each, and the loop is unrolled over the phases of
// both rings (lcm(kPF+2, 2) rows), so that slot indices are compile-time constants and
// advancing the march moves no registers (at the price of register pressure); otherwise
// the rolled loop shifts the ring by register moves.
template <typename TS, typename TF, typename M>
__device__ __forceinline__ uint64_t march2d(const typename Vec2<TS>::type* __restrict__ R2,
                                            const typename Vec2<TS>::type* __restrict__ M2,
                                            typename Vec2<TS>::type* __restrict__ RN2,
                                            typename Vec2<TS>::type* __restrict__ MN2, long long ys, int nb_x, int i0,
                                            int i1, int j, int jl, unsigned int ymask, bool owned, bool wrap_x, TF gamma,
                                            TF gmo, TF divgamma, TF K, TS DtDx, TS DtDy, uint64_t key)
{
#if HYDRO_2D_UNROLL
    using V2 = typename Vec2<TS>::type;
    constexpr int S = kPF + 2;                  // row slots: row r in slot (r - i0) % S
    constexpr int U = (S % 2 == 0) ? S : 2 * S;  // phases of the row ring and the 2-slot rings
    V2 ra, ma, qr[S], qm[S];
    load_cell<TS>(R2, M2, row_of(i0 - 1, nb_x, wrap_x) * ys + jl, ra, ma);
#pragma unroll
    for (int q = 0; q <= kPF; ++q)
        load_cell<TS>(R2, M2, row_of(min(i0 + q, i1 + 1), nb_x, wrap_x) * ys + jl, qr[q], qm[q]);
    Prim2<TF> P[2];  // X frame: un = u, ut = v; P[(i - i0) % 2] is row i
    TF F[2][4];      // F[(i - i0) % 2] is the X face (i-1, i)
    {
        Prim2<TF> Pm;
        Pm.r = static_cast<TF>(ra.x);
        cons_to_prim2<TF, M>(Pm.r, static_cast<TF>(ma.x), static_cast<TF>(ma.y), static_cast<TF>(ra.y), gmo, gamma,
                             Pm.un, Pm.ut, Pm.p, Pm.a);
        P[0].r = static_cast<TF>(qr[0].x);
        cons_to_prim2<TF, M>(P[0].r, static_cast<TF>(qm[0].x), static_cast<TF>(qm[0].y), static_cast<TF>(qr[0].y),
                             gmo, gamma, P[0].un, P[0].ut, P[0].p, P[0].a);
        face_flux2<TF, M>(Pm, P[0], K, divgamma, F[0][0], F[0][1], F[0][2], F[0][3]);
    }
    int i = i0;
    auto step = [&](auto phase) -> bool {
        constexpr int p = decltype(phase)::value;
        constexpr int sc = p % S, sn = (p + 1) % S, sl = (p + kPF + 1) % S;
        constexpr int pc = p % 2, pn = (p + 1) % 2;
        // Prefetch row i+1+kPF (clamped to the last row this strip reads) into the slot
        // row i-1 left.
        load_cell<TS>(R2, M2, row_of(min(i + 1 + kPF, i1 + 1), nb_x, wrap_x) * ys + jl, qr[sl], qm[sl]);
        key = row2d<TS, TF, M>(qr[sc], qm[sc], qr[sn], qm[sn], P[pc], P[pn], F[pc], F[pn], RN2, MN2, ys, i, j, ymask,
                               owned, gamma, gmo, divgamma, K, DtDx, DtDy, key);
        return ++i <= i1;
    };
    while (run_phases(step, std::make_integer_sequence<int, U>{})) {
    }
#else
    using V2 = typename Vec2<TS>::type;
    V2 ra, ma, rc, mc, qr[kPF + 1], qm[kPF + 1];
    load_cell<TS>(R2, M2, row_of(i0 - 1, nb_x, wrap_x) * ys + jl, ra, ma);
    load_cell<TS>(R2, M2, row_of(i0, nb_x, wrap_x) * ys + jl, rc, mc);
#pragma unroll
    for (int q = 0; q < kPF; ++q)
        load_cell<TS>(R2, M2, row_of(min(i0 + 1 + q, i1 + 1), nb_x, wrap_x) * ys + jl, qr[q], qm[q]);
    Prim2<TF> Pm, Pc;  // X frame: un = u, ut = v
    Pm.r = static_cast<TF>(ra.x);
    cons_to_prim2<TF, M>(Pm.r, static_cast<TF>(ma.x), static_cast<TF>(ma.y), static_cast<TF>(ra.y), gmo, gamma, Pm.un,
                         Pm.ut, Pm.p, Pm.a);
    Pc.r = static_cast<TF>(rc.x);
    cons_to_prim2<TF, M>(Pc.r, static_cast<TF>(mc.x), static_cast<TF>(mc.y), static_cast<TF>(rc.y), gmo, gamma, Pc.un,
                         Pc.ut, Pc.p, Pc.a);
    TF F1[4];
    face_flux2<TF, M>(Pm, Pc, K, divgamma, F1[0], F1[1], F1[2], F1[3]);
    for (int i = i0; i <= i1; ++i) {
        // Prefetch row i+1+kPF (clamped to the last row this strip reads).
        load_cell<TS>(R2, M2, row_of(min(i + 1 + kPF, i1 + 1), nb_x, wrap_x) * ys + jl, qr[kPF], qm[kPF]);
        Prim2<TF> Pn;
        TF F2[4];
        key = row2d<TS, TF, M>(rc, mc, qr[0], qm[0], Pc, Pn, F1, F2, RN2, MN2, ys, i, j, ymask, owned, gamma, gmo,
                               divgamma, K, DtDx, DtDy, key);
        // Advance the march.
        Pc = Pn;
        rc = qr[0];
        mc = qm[0];
#pragma unroll
        for (int q = 0; q < kPF; ++q) {
            qr[q] = qr[q + 1];
            qm[q] = qm[q + 1];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) F1[q] = F2[q];
    }
#endif
    return key;
}

}  // namespace hydro
