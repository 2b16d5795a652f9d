// The 3D block march shared by the single-domain step kernel and the batched kernel.
//
// A workgroup of NW = TY + 2 waves covers a (Y x Z) tile of TY rows x wz columns and
// marches along X over rows i0..i1 (reading i0-1 .. i1+1); wave w holds Y row
// j0 - 1 + w (waves 0 and TY+1 are halo rows: they load and convert their row but
// update nothing) and lane l holds Z column k:
//   * the Z face (k, k+1) is computed once by lane k and shuffled to lane k+1;
//   * the Y face (j, j+1) is computed once by the wave of row j from the primitives the
//     wave of row j+1 left in LDS, and left in LDS for the wave of row j+1 (two
//     workgroup barriers per X step, single-buffered LDS);
//   * the X face between rows i and i+1 is carried in registers along the march.
// Every wave of the workgroup must call march3d with the same i0, i1 (the barriers).
// The arithmetic of every cell is the reference's (hydro3d/kernel.cxx:145-407) in its
// operation order, so results do not depend on how the domain is tiled.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hydro_device.hpp"

namespace hydro {

constexpr int kOwnedZ = kWave - 2;  // most owned Z columns per window

// Z windows of a row of nb_z interior columns: the fewest windows of at most kOwnedZ
// columns, with the columns spread evenly over them, so that no window is a short
// straggler (512 columns: 9 windows of 57 rather than 8 of 62 and one of 16; on 512^3
// the even split makes the step's access pattern 5% faster).
__host__ __device__ __forceinline__ int z_windows(int nb_z) { return (nb_z + kOwnedZ - 1) / kOwnedZ; }
__host__ __device__ __forceinline__ int z_window_width(int nb_z)
{
    const int n = z_windows(nb_z);
    return (nb_z + n - 1) / n;
}
#ifndef HYDRO_3D_TY
#define HYDRO_3D_TY 6
#endif
// Owned Y rows per workgroup of the plain FP32-flux step: its FP32 arithmetic leaves the
// registers for 12-wave workgroups, one per CU, and the taller tile re-reads fewer halo
// rows (512^3: 2.33 vs 2.45 ms per step at 6 rows).  FP64 at 10 rows has too few waves
// per CU to hide its arithmetic (2.97 vs 2.67 ms).
#ifndef HYDRO_3D_TY_FLUX
#define HYDRO_3D_TY_FLUX 10
#endif
// Rows in flight ahead of row i+1 (register prefetch ring depth).
#ifndef HYDRO_3D_PREFETCH
#define HYDRO_3D_PREFETCH 1
#endif
// Minimum waves per SIMD of the 3D step kernel: 4 (two 8-wave workgroups per CU, at
// most 128 VGPRs) for FP64 state, 8 (four workgroups, at most 64 VGPRs) for FP32 state,
// whose step is latency-bound and runs 18% slower at two workgroups per CU than at three.
#ifndef HYDRO_3D_MIN_WAVES
#define HYDRO_3D_MIN_WAVES 4
#endif
#ifndef HYDRO_3D_MIN_WAVES_FP32
#define HYDRO_3D_MIN_WAVES_FP32 8
#endif
// Minimum waves per SIMD of the plain FP32-flux step: 3 (one 12-wave workgroup per CU).
// At 6 (two per CU, within 80 VGPRs with no load ahead) it ran 5% slower on 512^3.
#ifndef HYDRO_3D_MIN_WAVES_FLUX
#define HYDRO_3D_MIN_WAVES_FLUX 3
#endif
constexpr int kTY = HYDRO_3D_TY;           // owned Y rows per workgroup (FP64, FP32 state, batched)
constexpr int kTYFlux = HYDRO_3D_TY_FLUX;  // owned Y rows per workgroup of the plain FP32-flux step
constexpr int kPF3 = HYDRO_3D_PREFETCH;
constexpr int kNW = kTY + 2;               // waves per workgroup at kTY
constexpr int kBlock3 = kNW * kWave;

template <typename T>
struct Vec2_3;
template <>
struct Vec2_3<double> {
    using type = double2;
};
template <>
struct Vec2_3<float> {
    using type = float2;
};

template <typename T>
struct Trip {
    T x, y, z;
};

template <typename T>
struct Consts3 {
    T C_min, gamma, gmo, divgamma, K;
};

__device__ __forceinline__ unsigned xcd_remap3(unsigned b, unsigned n)
{
    constexpr unsigned X = 8;
    const unsigned q = n / X, r = n % X;
    const unsigned xcd = b % X, k = b / X;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// Tile row held by the calling wave (0 and NW-1 are the halo rows).  The waves of a
// workgroup sit two or more per SIMD, and a halo wave does a fraction of an owned row's work.
// The halo rows go to waves on SIMDs 0 and 1 in even workgroup slots of the CU and on
// SIMDs 2 and 3 in odd ones (HW_ID register), so that co-resident workgroups spread the
// owned rows evenly over the four SIMDs.  Any outcome is a permutation of the rows, and
// the march's results do not depend on which wave holds which row.  Every thread of
// the workgroup must call it (one barrier).
template <int NW>
__device__ __forceinline__ int tile_row_of_wave(int (&s_simd)[NW])
{
    const int hw = static_cast<int>(threadIdx.x) / kWave;
    const unsigned id = __builtin_amdgcn_s_getreg((31 << 11) | 4);  // HW_REG_HW_ID, bits 31:0
    const int simd = static_cast<int>((id >> 4) & 3u), slot = static_cast<int>((id >> 16) & 15u);
    if ((threadIdx.x & (kWave - 1)) == 0) s_simd[hw] = simd;
    __syncthreads();
    const int want0 = (slot & 1) * 2, want1 = want0 + 1;
    int lo = 0;  // first wave on SIMD want0 (wave 0 if none)
    for (int v = NW - 1; v >= 0; --v)
        if (s_simd[v] == want0) lo = v;
    int hi = lo == NW - 1 ? NW - 2 : NW - 1;  // first other wave on SIMD want1
    for (int v = NW - 1; v >= 0; --v)
        if (v != lo && s_simd[v] == want1) hi = v;
    if (hw == lo) return 0;
    if (hw == hi) return NW - 1;
    return 1 + hw - (lo < hw ? 1 : 0) - (hi < hw ? 1 : 0);
}

// Primitive state in the X frame: (rho, u, v, w, p, a).
template <typename T>
struct P6 {
    T r, u, v, w, p, a;
};

template <typename T>
__device__ __forceinline__ void load3(const typename Vec2_3<T>::type* __restrict__ R2,
                                      const Trip<T>* __restrict__ M3, long long n, typename Vec2_3<T>::type& re,
                                      Trip<T>& m)
{
    re = R2[n];
    m = M3[n];
}

template <typename TS, typename TF, typename M>
__device__ __forceinline__ void to_prim(const typename Vec2_3<TS>::type& re, const Trip<TS>& m, TF gmo, TF gamma,
                                        P6<TF>& q)
{
    q.r = static_cast<TF>(re.x);
    cons_to_prim3<TF, M>(q.r, static_cast<TF>(m.x), static_cast<TF>(m.y), static_cast<TF>(m.z), static_cast<TF>(re.y),
                         gmo, gamma, q.u, q.v, q.w, q.p, q.a);
}

template <typename T>
__device__ __forceinline__ Prim3<T> frame_x(const P6<T>& q)
{
    return Prim3<T>{ q.r, q.u, q.v, q.w, q.p, q.a };
}
template <typename T>
__device__ __forceinline__ Prim3<T> frame_y(const P6<T>& q)
{
    return Prim3<T>{ q.r, q.v, q.u, q.w, q.p, q.a };  // hydro3d/kernel.cxx:254-256
}
template <typename T>
__device__ __forceinline__ Prim3<T> frame_z(const P6<T>& q)
{
    return Prim3<T>{ q.r, q.w, q.v, q.u, q.p, q.a };  // hydro3d/kernel.cxx:296-298
}

// Row actually read for row r (see row_of in the 2D march): wrap_x reads the periodic X
// ghost planes from the opposite interior planes.
__device__ __forceinline__ long long row3_of(int r, int nb_x, bool wrap_x)
{
    return wrap_x ? (r == 0 ? nb_x : (r == nb_x + 1 ? 1 : r)) : r;
}

template <typename TF, int NW>
struct March3Lds {
    TF prim[NW][6][kWave];  // Y-frame primitives (rho, v, u, w, p, a) of every wave
    TF flux[NW][5][kWave];  // Y-face fluxes of every wave
};

// March rows i0..i1 of one (Y x Z) tile; see the file comment.  `col` is the in-plane
// offset jl * zs + kl of the column this lane loads; stores go to the same column.  PF
// rows of loads run ahead of row i+1 (0: each row's loads are issued where it is used).
template <typename TS, typename TF, typename M, int PF, int NW>
__device__ __forceinline__ uint64_t march3d(const typename Vec2_3<TS>::type* __restrict__ R2,
                                            const Trip<TS>* __restrict__ M3,
                                            typename Vec2_3<TS>::type* __restrict__ RN2,
                                            Trip<TS>* __restrict__ MN3, long long zy, long long col, int nb_x, int i0,
                                            int i1, int w, int lane, bool halo_wave, bool owned, unsigned int ymask,
                                            unsigned int zmask, bool wrap_x, TF gamma, TF gmo, TF divgamma, TF K,
                                            TS DtDx, TS DtDy, TS DtDz, March3Lds<TF, NW>& lds, uint64_t key)
{
    using V2 = typename Vec2_3<TS>::type;
    auto& s_prim = lds.prim;
    auto& s_flux = lds.flux;
    // Row i in rc/mc; rows i+1 .. i+PF in the ring qr/qm (slot PF takes the next
    // load each step).
    V2 rc, qr[PF + 1];
    Trip<TS> mc, qm[PF + 1];
    P6<TF> Pc;
    TF fx1_0 = TF(0), fx1_1 = TF(0), fx1_2 = TF(0), fx1_3 = TF(0), fx1_4 = TF(0);
    load3<TS>(R2, M3, row3_of(i0, nb_x, wrap_x) * zy + col, rc, mc);
#pragma unroll
    for (int q = 0; q < PF; ++q)
        load3<TS>(R2, M3, row3_of(min(i0 + 1 + q, i1 + 1), nb_x, wrap_x) * zy + col, qr[q], qm[q]);
    to_prim<TS, TF, M>(rc, mc, gmo, gamma, Pc);
    if (!halo_wave) {
        V2 ra;
        Trip<TS> ma;
        P6<TF> Pm;
        load3<TS>(R2, M3, row3_of(i0 - 1, nb_x, wrap_x) * zy + col, ra, ma);
        to_prim<TS, TF, M>(ra, ma, gmo, gamma, Pm);
        face_flux3<TF, M>(frame_x(Pm), frame_x(Pc), K, divgamma, fx1_0, fx1_1, fx1_2, fx1_3, fx1_4);
    }

    for (int i = i0; i <= i1; ++i) {
        const int ipf = min(i + 1 + PF, i1 + 1);
        load3<TS>(R2, M3, row3_of(ipf, nb_x, wrap_x) * zy + col, qr[PF], qm[PF]);
        const V2 rn = qr[0];
        const Trip<TS> mn = qm[0];

        // (a) publish this row's Y-frame primitives.
        const Prim3<TF> Sy = frame_y(Pc);
        s_prim[w][0][lane] = Sy.r;
        s_prim[w][1][lane] = xor_sign(Sy.un, ymask);
        s_prim[w][2][lane] = Sy.ut1;
        s_prim[w][3][lane] = Sy.ut2;
        s_prim[w][4][lane] = Sy.p;
        s_prim[w][5][lane] = Sy.a;
        __syncthreads();

        // (b) Y face (j, j+1), hydro3d/kernel.cxx:239-278, evaluated by the wave of row j
        // from the primitives the wave of row j+1 published.
        if (w < NW - 1) {
            Prim3<TF> L, R;
            L.r = s_prim[w][0][lane];
            L.un = s_prim[w][1][lane];
            L.ut1 = s_prim[w][2][lane];
            L.ut2 = s_prim[w][3][lane];
            L.p = s_prim[w][4][lane];
            L.a = s_prim[w][5][lane];
            R.r = s_prim[w + 1][0][lane];
            R.un = s_prim[w + 1][1][lane];
            R.ut1 = s_prim[w + 1][2][lane];
            R.ut2 = s_prim[w + 1][3][lane];
            R.p = s_prim[w + 1][4][lane];
            R.a = s_prim[w + 1][5][lane];
            TF f0, f1, f2, f3, f4;
            face_flux3<TF, M>(L, R, K, divgamma, f0, f1, f2, f3, f4);
            s_flux[w][0][lane] = f0;
            s_flux[w][1][lane] = f1;
            s_flux[w][2][lane] = f2;
            s_flux[w][3][lane] = f3;
            s_flux[w][4][lane] = f4;
        }
        __syncthreads();

        V2 rcur = rc;
        Trip<TS> mcur = mc;
        P6<TF> Pn;
        to_prim<TS, TF, M>(rn, mn, gmo, gamma, Pn);
        if (!halo_wave) {
            // (c) Y face above this row, from the wave of row j-1.
            const TF fy1_0 = s_flux[w - 1][0][lane];
            const TF fy1_1 = s_flux[w - 1][1][lane];
            const TF fy1_2 = s_flux[w - 1][2][lane];
            const TF fy1_3 = s_flux[w - 1][3][lane];
            const TF fy1_4 = s_flux[w - 1][4][lane];
            const TF fy2_0 = s_flux[w][0][lane];
            const TF fy2_1 = s_flux[w][1][lane];
            const TF fy2_2 = s_flux[w][2][lane];
            const TF fy2_3 = s_flux[w][3][lane];
            const TF fy2_4 = s_flux[w][4][lane];

            // Z face (k, k+1) by lane k, shuffled to lane k+1, hydro3d/kernel.cxx:281-320.
            Prim3<TF> Sz = frame_z(Pc);
            Sz.un = xor_sign(Sz.un, zmask);
            Prim3<TF> Nz;
            Nz.r = from_next_lane(Sz.r);
            Nz.un = from_next_lane(Sz.un);
            Nz.ut1 = from_next_lane(Sz.ut1);
            Nz.ut2 = from_next_lane(Sz.ut2);
            Nz.p = from_next_lane(Sz.p);
            Nz.a = from_next_lane(Sz.a);
            TF fz2_0, fz2_1, fz2_2, fz2_3, fz2_4;
            face_flux3<TF, M>(Sz, Nz, K, divgamma, fz2_0, fz2_1, fz2_2, fz2_3, fz2_4);
            const TF fz1_0 = from_prev_lane(fz2_0);
            const TF fz1_1 = from_prev_lane(fz2_1);
            const TF fz1_2 = from_prev_lane(fz2_2);
            const TF fz1_3 = from_prev_lane(fz2_3);
            const TF fz1_4 = from_prev_lane(fz2_4);

            // X face (i, i+1).
            TF fx2_0, fx2_1, fx2_2, fx2_3, fx2_4;
            face_flux3<TF, M>(frame_x(Pc), frame_x(Pn), K, divgamma, fx2_0, fx2_1, fx2_2, fx2_3, fx2_4);

            // Update, hydro3d/kernel.cxx:330-358: U' = ((U + x) + y) + z, in TS.
            const TS x_r = DtDx * (static_cast<TS>(fx1_0) - static_cast<TS>(fx2_0));
            const TS x_u = DtDx * (static_cast<TS>(fx1_1) - static_cast<TS>(fx2_1));
            const TS x_v = DtDx * (static_cast<TS>(fx1_2) - static_cast<TS>(fx2_2));
            const TS x_w = DtDx * (static_cast<TS>(fx1_3) - static_cast<TS>(fx2_3));
            const TS x_E = DtDx * (static_cast<TS>(fx1_4) - static_cast<TS>(fx2_4));
            const TS y_r = DtDy * (static_cast<TS>(fy1_0) - static_cast<TS>(fy2_0));
            const TS y_u = DtDy * (static_cast<TS>(fy1_2) - static_cast<TS>(fy2_2));
            const TS y_v = DtDy * (static_cast<TS>(fy1_1) - static_cast<TS>(fy2_1));
            const TS y_w = DtDy * (static_cast<TS>(fy1_3) - static_cast<TS>(fy2_3));
            const TS y_E = DtDy * (static_cast<TS>(fy1_4) - static_cast<TS>(fy2_4));
            const TS z_r = DtDz * (static_cast<TS>(fz1_0) - static_cast<TS>(fz2_0));
            const TS z_u = DtDz * (static_cast<TS>(fz1_3) - static_cast<TS>(fz2_3));
            const TS z_v = DtDz * (static_cast<TS>(fz1_2) - static_cast<TS>(fz2_2));
            const TS z_w = DtDz * (static_cast<TS>(fz1_1) - static_cast<TS>(fz2_1));
            const TS z_E = DtDz * (static_cast<TS>(fz1_4) - static_cast<TS>(fz2_4));
            const TS nr = rcur.x + x_r + y_r + z_r;
            const TS nu = mcur.x + x_u + y_u + z_u;
            const TS nv = mcur.y + x_v + y_v + z_v;
            const TS nw = mcur.z + x_w + y_w + z_w;
            const TS nE = rcur.y + x_E + y_E + z_E;
            const TF sp = signal_speed3<TF, M>(static_cast<TF>(nr), static_cast<TF>(nu), static_cast<TF>(nv),
                                               static_cast<TF>(nw), static_cast<TF>(nE), gmo, gamma);
            const uint64_t k_new = speed_key(static_cast<double>(sp));
            key = owned ? umax64(key, k_new) : key;
            if (owned) {
                const long long n = static_cast<long long>(i) * zy + col;
                V2 o;
                o.x = nr;
                o.y = nE;
                RN2[n] = o;
                Trip<TS> m;
                m.x = nu;
                m.y = nv;
                m.z = nw;
                MN3[n] = m;
            }
            fx1_0 = fx2_0;
            fx1_1 = fx2_1;
            fx1_2 = fx2_2;
            fx1_3 = fx2_3;
            fx1_4 = fx2_4;
        }
        Pc = Pn;
        rc = rn;
        mc = mn;
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            qr[q] = qr[q + 1];
            qm[q] = qm[q + 1];
        }
    }

    return key;
}

}  // namespace hydro
