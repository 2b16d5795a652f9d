// The 2D strip march shared by the single-domain step kernel and the batched kernel.
//
// One wave (64 lanes, lane l <-> Y column j) marches along X over rows i0..i1 of its
// column window, from i0 up or from i1 down, reading rows i0-1 .. i1+1:
//   * the X face between rows i and i+1 is computed once and carried in registers to
//     the next row (fx2 of row i is fx1 of row i+1; hydro2d/kernel.cxx:186-210 computes
//     it twice with identical arguments);
//   * the Y face (j, j+1) is computed once by lane j and handed to lane j+1 by a
//     cross-lane shuffle (hydro2d/kernel.cxx:213-251);
//   * primitives and the impedance sqrt(gamma*p*rho) are computed once per cell;
//   * kPF rows of loads are kept in flight ahead of the row being converted.
// The arithmetic of every cell is the reference's (hydro2d/kernel.cxx:132-319) in its
// operation order, so results do not depend on how rows and windows are split or on
// the direction of the march.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hydro_device.hpp"

namespace hydro {

constexpr int kOwned = kWave - 2;  // owned columns per wave window
// Waves (adjacent column windows) per workgroup: 4, a 4 KiB row segment per array.
// Workgroups of 2 waves at the same waves per CU ran the FP64 step 6% slower, of 8
// waves 3% slower (40000x20000, alternated in one process).
#ifndef HYDRO_2D_WAVES
#define HYDRO_2D_WAVES 4
#endif
constexpr int kWavesPerBlock = HYDRO_2D_WAVES;
// Workgroups per CU the 2D step must fit in registers (the launch bound; the launch
// itself runs fewer, see default_occ_2d): 5 (at most 102 VGPRs), and 4 for the FP64 scheme
// (at most 128; it uses 90 with one row of loads ahead).
#ifndef HYDRO_2D_MIN_BLOCKS
#define HYDRO_2D_MIN_BLOCKS 5
#endif
#ifndef HYDRO_2D_MIN_BLOCKS_FP64
#define HYDRO_2D_MIN_BLOCKS_FP64 4
#endif
// Rows of loads issued ahead of the row being converted (register prefetch ring depth).
// The ring advances by register moves, before which the compiler waits for the loads in
// flight, so about one row is in flight whatever the depth; depth 1 needs the fewest
// registers and moves (FP64 90 VGPRs) and ran 0.3% (FP64), 2.2% (FP32-flux) and 0.7% (FP32
// state) faster than 3 on 40000x20000, alternated in one process.
#ifndef HYDRO_2D_PREFETCH
#define HYDRO_2D_PREFETCH 1
#endif
constexpr int kPF = HYDRO_2D_PREFETCH;
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
// of the next row n (into Pn) and the X face between rows i and n (into F2), the update
// of row i with the X face carried from the previous row in F1, its store and the order
// key of its new signal speed.  Forward (REV false): n = i+1, F1 is the face (i-1, i)
// and F2 the face (i, i+1).  Reverse: n = i-1, F1 is the face (i, i+1) and F2 the face
// (i-1, i).  Either way a face is evaluated as flux(lower row, upper row) and the update
// takes (lower face - upper face), so both directions give the same bits.
template <typename TS, typename TF, typename M, bool REV>
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

    // Next row's primitives and the X face between rows i and n.
    Pn.r = static_cast<TF>(rn.x);
    cons_to_prim2<TF, M>(Pn.r, static_cast<TF>(mn.x), static_cast<TF>(mn.y), static_cast<TF>(rn.y), gmo, gamma, Pn.un,
                         Pn.ut, Pn.p, Pn.a);
    if constexpr (REV) face_flux2<TF, M>(Pn, Pc, K, divgamma, F2[0], F2[1], F2[2], F2[3]);
    else face_flux2<TF, M>(Pc, Pn, K, divgamma, F2[0], F2[1], F2[2], F2[3]);
    const TF* FL = REV ? F2 : F1;  // face (i-1, i)
    const TF* FU = REV ? F1 : F2;  // face (i, i+1)

    // Update, kernel.cxx:259-278: U' = (U + x) + y; Y fluxes are in the (rho, v, u, E)
    // frame, so rho*u takes component 2 and rho*v component 1.  Increments and update
    // are in the state precision TS.
    const TS x_r = DtDx * (static_cast<TS>(FL[0]) - static_cast<TS>(FU[0]));
    const TS x_u = DtDx * (static_cast<TS>(FL[1]) - static_cast<TS>(FU[1]));
    const TS x_v = DtDx * (static_cast<TS>(FL[2]) - static_cast<TS>(FU[2]));
    const TS x_E = DtDx * (static_cast<TS>(FL[3]) - static_cast<TS>(FU[3]));
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

// March rows i0..i1 (i0 <= i1) of one window, from i0 up (REV false) or from i1 down.
// Lane columns: j (global, for stores and ownership) and jl (the column loaded); wrap_x
// as in row_of.  Returns `key` folded with the order keys of the new signal speeds of
// the owned cells.  PF rows of loads are in flight while a row is computed.
//
// The batched form's direct dispatch alternates the direction by tile row, so that two
// vertically adjacent tiles running side by side read their shared halo rows at the
// same moment (both at their start or both at their end) rather than a tile's length
// apart: 1.6-2% faster on the 256-blast mix.  The plain step marches forward only:
// alternating its strips made it 1-4% slower (its halo rows already hit in cache, and
// strips marching apart widen the band of rows in flight).
template <typename TS, typename TF, typename M, bool REV, int PF = kPF>
__device__ __forceinline__ uint64_t march2d(const typename Vec2<TS>::type* __restrict__ R2,
                                            const typename Vec2<TS>::type* __restrict__ M2,
                                            typename Vec2<TS>::type* __restrict__ RN2,
                                            typename Vec2<TS>::type* __restrict__ MN2, long long ys, int nb_x, int i0,
                                            int i1, int j, int jl, unsigned int ymask, bool owned, bool wrap_x, TF gamma,
                                            TF gmo, TF divgamma, TF K, TS DtDx, TS DtDy, uint64_t key)
{
    using V2 = typename Vec2<TS>::type;
    constexpr int d = REV ? -1 : 1;
    const int first = REV ? i1 : i0, last = REV ? i0 : i1;
    // The rows read are i0-1 .. i1+1; `clamp` keeps the prefetch inside them.
    auto clamp = [&](int r) { return REV ? max(r, i0 - 1) : min(r, i1 + 1); };
    V2 ra, ma, rc, mc, qr[PF + 1], qm[PF + 1];
    load_cell<TS>(R2, M2, row_of(first - d, nb_x, wrap_x) * ys + jl, ra, ma);
    load_cell<TS>(R2, M2, row_of(first, nb_x, wrap_x) * ys + jl, rc, mc);
#pragma unroll
    for (int q = 0; q < PF; ++q)
        load_cell<TS>(R2, M2, row_of(clamp(first + d * (1 + q)), nb_x, wrap_x) * ys + jl, qr[q], qm[q]);
    Prim2<TF> Pm, Pc;  // X frame: un = u, ut = v
    Pm.r = static_cast<TF>(ra.x);
    cons_to_prim2<TF, M>(Pm.r, static_cast<TF>(ma.x), static_cast<TF>(ma.y), static_cast<TF>(ra.y), gmo, gamma, Pm.un,
                         Pm.ut, Pm.p, Pm.a);
    Pc.r = static_cast<TF>(rc.x);
    cons_to_prim2<TF, M>(Pc.r, static_cast<TF>(mc.x), static_cast<TF>(mc.y), static_cast<TF>(rc.y), gmo, gamma, Pc.un,
                         Pc.ut, Pc.p, Pc.a);
    TF F1[4];  // the face between the first row and the row behind it
    if constexpr (REV) face_flux2<TF, M>(Pc, Pm, K, divgamma, F1[0], F1[1], F1[2], F1[3]);
    else face_flux2<TF, M>(Pm, Pc, K, divgamma, F1[0], F1[1], F1[2], F1[3]);
    for (int i = first; REV ? i >= last : i <= last; i += d) {
        // Prefetch the row PF+1 ahead (clamped to the last row this strip reads).
        load_cell<TS>(R2, M2, row_of(clamp(i + d * (1 + PF)), nb_x, wrap_x) * ys + jl, qr[PF], qm[PF]);
        Prim2<TF> Pn;
        TF F2[4];
        key = row2d<TS, TF, M, REV>(rc, mc, qr[0], qm[0], Pc, Pn, F1, F2, RN2, MN2, ys, i, j, ymask, owned, gamma, gmo,
                                    divgamma, K, DtDx, DtDy, key);
        // Advance the march.
        Pc = Pn;
        rc = qr[0];
        mc = qm[0];
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            qr[q] = qr[q + 1];
            qm[q] = qm[q + 1];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) F1[q] = F2[q];
    }
    return key;
}

// march2d in the direction `rev` (wave-uniform).
template <typename TS, typename TF, typename M, int PF = kPF>
__device__ __forceinline__ uint64_t march2d_dir(bool rev, const typename Vec2<TS>::type* __restrict__ R2,
                                                const typename Vec2<TS>::type* __restrict__ M2,
                                                typename Vec2<TS>::type* __restrict__ RN2,
                                                typename Vec2<TS>::type* __restrict__ MN2, long long ys, int nb_x,
                                                int i0, int i1, int j, int jl, unsigned int ymask, bool owned,
                                                bool wrap_x, TF gamma, TF gmo, TF divgamma, TF K, TS DtDx, TS DtDy,
                                                uint64_t key)
{
    if (rev)
        return march2d<TS, TF, M, true, PF>(R2, M2, RN2, MN2, ys, nb_x, i0, i1, j, jl, ymask, owned, wrap_x, gamma, gmo,
                                        divgamma, K, DtDx, DtDy, key);
    return march2d<TS, TF, M, false, PF>(R2, M2, RN2, MN2, ys, nb_x, i0, i1, j, jl, ymask, owned, wrap_x, gamma, gmo,
                                     divgamma, K, DtDx, DtDy, key);
}

}  // namespace hydro
