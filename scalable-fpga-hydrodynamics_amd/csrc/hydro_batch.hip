// Batched form: many independent blasts of uneven sizes advanced together, one
// persistent kernel per step (plus a one-block-per-blast finalise), optionally replayed
// from a hipGraph.
//
// Each blast is a complete solve of the reference scheme (hydro2d/kernel.cxx:132-319,
// hydro3d/kernel.cxx:145-407, step loop hydro2d/main.cxx:686-815) with the reference's
// boundaries: X periodic, Y (and Z) reflective.  Its state stays in its own two buffer
// pairs in the reference layout; its Dt, t, iteration count, stop flag and current
// buffer live in a device record (BatchBlast), so steps need no host round trip and
// blasts finish independently.
//
// Work decomposition.  The interior of every blast is cut into "columns": a 2D column is
// a 62-wide Y window over all X rows (one wave marches it), a 3D column a 6-row Y tile of
// a Z window over all X rows (one 8-wave workgroup marches it; the Z windows of a row
// share its columns evenly, at most 62 each).  The persistent grid has a
// fixed number of workers (waves in 2D, workgroups in 3D); a host-built segment table
// gives each worker a list of (blast, column, rows i0..i1) segments.  Two tilings:
//   balanced   the concatenated rows of all columns are split into equal contiguous
//              ranges, one per worker (equal work, long marches, few halo rows);
//   R-row      columns are cut into tiles of R rows, in row-major order within each
//              blast, dealt round-robin to the workers; 2D tiles of one window are
//              instead dispatched directly, one wave per tile (see direct_mode).
// The autotuner times the candidates (grid occupancy x tiling) on the batch itself.
// Since every cell's arithmetic is the reference's whatever the tiling, tuning steps
// are ordinary steps of the solve.
//
// Periodic X is read directly from the opposite interior plane (march WRAP_X), so the
// batched form needs no ghost-fill launches; Y/Z walls are the branch-free mirrors of
// the single-domain kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <new>
#include <string.h>
#include <string>
#include <vector>

#include "../../include/hydro_hip.h"
#include "hydro_device.hpp"
#include "hydro_kernels.hpp"
#include "hydro_march2d.hpp"
#include "hydro_march3d.hpp"
#include "hydro_runtime.hpp"

namespace hydro {

struct BatchBlast {
    void* rhoE[2];  // buffer pairs; `cur` holds the current state
    void* mom[2];
    int nb_x, nb_y, nb_z;
    int active;
    long long ys, zs, zy;  // 2D: ys = nb_y + 2; 3D: zs = nb_z + 2, zy = (nb_y + 2) * zs
    unsigned int cur;
    unsigned int pad;
    unsigned long long ic_flags;  // states present after a device IC (bit 0 in, bit 1 out)
    double dt_in, dt_out;         // Dt of the two IC states
    StepControl ctl;
};

struct BatchSeg {
    int blast;
    int col;  // 2D: Y window; 3D: yt * nzw + zw
    int i0, i1;
};

namespace {

// Waves per SIMD of the 2D batch kernel (= workgroups per CU): four, at most 128 VGPRs
// (FP64 uses 128, FP32-flux 96; FP32-flux at five would spill).
// Rows of loads in flight in the batched march: 2, one fewer than the plain step's.
// Its tiles are 16 rows long; on the 256-blast mix 1 or 2 rows ran 0.5% faster than 3,
// and 4 or 5 slower.
#ifndef HYDRO_BATCH_PREFETCH
#define HYDRO_BATCH_PREFETCH 2
#endif
constexpr int kPFBatch = HYDRO_BATCH_PREFETCH;
template <typename TF>
constexpr int kBatchWaves2 = 4;

__device__ __forceinline__ void flush_key(BatchBlast* __restrict__ blasts, int b, uint64_t key, unsigned wid, int lane)
{
    const uint64_t k = wave_max_u64(key);
    if (lane == 0 && k != 0) atomicMax(&blasts[b].ctl.red.slot[wid % kRedSlots], static_cast<unsigned long long>(k));
}

template <typename TS, typename TF, typename M>
__global__ void __launch_bounds__(kBlock, kBatchWaves2<TF>)
    batch2d_kernel(BatchBlast* __restrict__ blasts, const BatchSeg* __restrict__ segs,
                   const unsigned* __restrict__ seg_off, Consts<double> cs, unsigned nworkers, int group, int direct)
{
    using V2 = typename Vec2<TS>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    // A worker is a group of `group` waves of one workgroup marching adjacent windows
    // (col * group + g) over the same rows, so that a row of the group is one contiguous
    // stream of group KiB per array.  Wave-uniform values are made explicit so that the
    // segment and record loads are scalar.
    const int wv = threadIdx.x / kWave;
    const int g = __builtin_amdgcn_readfirstlane(wv % group);
    const unsigned wid = __builtin_amdgcn_readfirstlane(lb * (kWavesPerBlock / group) + wv / group);
    if (wid >= nworkers) return;  // never in direct mode (its only barrier is at the end)
    const TF gamma = static_cast<TF>(cs.gamma), gmo = static_cast<TF>(cs.gmo);
    const TF divgamma = static_cast<TF>(cs.divgamma), K = static_cast<TF>(cs.K);
    // Direct dispatch: wave `wid` takes segment `wid` of one table.
    const unsigned s0 = direct ? wid : seg_off[wid];
    const unsigned s1 = direct ? (wid < seg_off[1] ? wid + 1 : wid) : seg_off[wid + 1];
    int cb = -1;
    uint64_t key = 0;
    for (unsigned s = s0; s < s1; ++s) {
        const BatchSeg sg = segs[s];
        if (sg.blast != cb) {
            if (cb >= 0) flush_key(blasts, cb, key, wid, lane);
            key = 0;
            cb = sg.blast;
        }
        const BatchBlast* B = blasts + sg.blast;
        if (B->ctl.done) continue;
        const unsigned cur = B->cur;
        const TS dt = static_cast<TS>(B->ctl.dt);
        const TS DtDx = dt / static_cast<TS>(B->ctl.D[0]);  // hydro2d/main.cxx:703-704
        const TS DtDy = dt / static_cast<TS>(B->ctl.D[1]);
        const int nb_x = B->nb_x, nb_y = B->nb_y;
        const long long ys = B->ys;
        const V2* R2 = static_cast<const V2*>(B->rhoE[cur]);
        const V2* M2 = static_cast<const V2*>(B->mom[cur]);
        V2* RN2 = static_cast<V2*>(B->rhoE[cur ^ 1u]);
        V2* MN2 = static_cast<V2*>(B->mom[cur ^ 1u]);
        const int win = sg.col * group + g;
        if (win * kOwned >= nb_y) continue;  // past the blast's last window
        const int j = win * kOwned + lane;
        const bool owned = lane >= 1 && lane <= kOwned && j <= nb_y;
        int jl;
        unsigned int ymask;
        window_lane2d<TF, true>(j, nb_y, false, jl, ymask);
        // A segment with i0 > i1 marches its rows i1..i0 downward (odd tile rows of direct
        // dispatch, so that vertically adjacent tiles read their shared halo rows together).
        const bool rev = sg.i0 > sg.i1;
        key = march2d_dir<TS, TF, M, kPFBatch>(rev, R2, M2, RN2, MN2, ys, nb_x, rev ? sg.i1 : sg.i0,
                                               rev ? sg.i0 : sg.i1, j, jl, ymask, owned, true, gamma, gmo, divgamma,
                                               K, DtDx, DtDy, key);
    }
    if (direct) {
        // Every wave of a direct workgroup reaches this point.  Its waves hold consecutive
        // tiles, mostly of one blast: fold their keys so that the workgroup issues one
        // atomic per blast, as the plain step does (one per wave made 60% of the step on
        // one large blast with 12-row tiles, serialised on the blast's 64 slots).
        __shared__ unsigned long long s_key[kWavesPerBlock];
        __shared__ int s_blast[kWavesPerBlock];
        const uint64_t k = wave_max_u64(key);
        if (lane == 0) {
            s_key[wv] = k;
            s_blast[wv] = cb;
        }
        __syncthreads();
        if (threadIdx.x < kWavesPerBlock) {
            const int t = threadIdx.x;
            const int bl = s_blast[t];
            if (bl >= 0 && (t == 0 || s_blast[t - 1] != bl)) {
                uint64_t m = s_key[t];
                for (int u = t + 1; u < kWavesPerBlock && s_blast[u] == bl; ++u) m = umax64(m, s_key[u]);
                if (m != 0)
                    atomicMax(&blasts[bl].ctl.red.slot[blockIdx.x % kRedSlots], static_cast<unsigned long long>(m));
            }
        }
        return;
    }
    if (cb >= 0) flush_key(blasts, cb, key, wid, lane);
}

// Four waves per SIMD (two 8-wave workgroups per CU): at most 128 VGPRs.
template <typename TS, typename TF, typename M>
__global__ void __launch_bounds__(kBlock3, 4)
    batch3d_kernel(BatchBlast* __restrict__ blasts, const BatchSeg* __restrict__ segs,
                   const unsigned* __restrict__ seg_off, Consts<double> cs, unsigned nworkers)
{
    using V2 = typename Vec2_3<TS>::type;
    __shared__ March3Lds<TF, kNW> lds;
    const unsigned wid = xcd_remap3(blockIdx.x, gridDim.x);
    if (wid >= nworkers) return;  // workgroup-uniform
    const int lane = threadIdx.x & (kWave - 1);
    __shared__ int s_simd[kNW];
    const int w = tile_row_of_wave(s_simd);  // tile row of this wave (halo rows placed by SIMD)
    const bool halo_wave = (w == 0) || (w == kNW - 1);
    const TF gamma = static_cast<TF>(cs.gamma), gmo = static_cast<TF>(cs.gmo);
    const TF divgamma = static_cast<TF>(cs.divgamma), K = static_cast<TF>(cs.K);
    const unsigned s0 = seg_off[wid], s1 = seg_off[wid + 1];
    int cb = -1;
    uint64_t key = 0;
    for (unsigned s = s0; s < s1; ++s) {
        const BatchSeg sg = segs[s];
        if (sg.blast != cb) {
            if (cb >= 0) flush_key(blasts, cb, key, wid, lane);
            key = 0;
            cb = sg.blast;
        }
        const BatchBlast* B = blasts + sg.blast;
        if (B->ctl.done) continue;  // workgroup-uniform: every wave reads the same record
        const unsigned cur = B->cur;
        const TS dt = static_cast<TS>(B->ctl.dt);
        const TS DtDx = dt / static_cast<TS>(B->ctl.D[0]);  // hydro3d/main.cxx:759-761
        const TS DtDy = dt / static_cast<TS>(B->ctl.D[1]);
        const TS DtDz = dt / static_cast<TS>(B->ctl.D[2]);
        const int nb_x = B->nb_x, nb_y = B->nb_y, nb_z = B->nb_z;
        const int nzw = z_windows(nb_z), wz = z_window_width(nb_z);
        const int yt = sg.col / nzw, zw = sg.col % nzw;
        const int j = yt * kTY + w;
        const int k = zw * wz + lane;
        const bool owned = !halo_wave && j <= nb_y && lane >= 1 && lane <= wz && k <= nb_z;
        const int jl = j < 1 ? 1 : (j > nb_y ? nb_y : j);
        const unsigned int ymask = (j == 0 || j == nb_y + 1) ? kSignMask<TF> : 0u;
        const int kl = k < 1 ? 1 : (k > nb_z ? nb_z : k);
        const unsigned int zmask = (k == 0 || k == nb_z + 1) ? kSignMask<TF> : 0u;
        const long long col = static_cast<long long>(jl) * B->zs + kl;
        key = march3d<TS, TF, M, kPF3>(static_cast<const V2*>(B->rhoE[cur]), static_cast<const Trip<TS>*>(B->mom[cur]),
                                 static_cast<V2*>(B->rhoE[cur ^ 1u]), static_cast<Trip<TS>*>(B->mom[cur ^ 1u]), B->zy,
                                 col, nb_x, sg.i0, sg.i1, w, lane, halo_wave, owned, ymask, zmask, true, gamma, gmo,
                                 divgamma, K, DtDx, DtDy, DtDz, lds, key);
    }
    if (cb >= 0) flush_key(blasts, cb, key, wid, lane);
}

// One workgroup (one wave) per slot: fold the slots into Dt = C_min / max speed
// (hydro2d/kernel.cxx:318), advance t += Dt and it (hydro2d/main.cxx:813-814), stop on
// !(t < max_t && it < max_it), flip the current buffer and publish the stop flag.
__global__ void __launch_bounds__(kWave) batch_finalize_kernel(BatchBlast* __restrict__ blasts,
                                                               unsigned char* __restrict__ done_out, unsigned nslots)
{
    const unsigned b = blockIdx.x;
    if (b >= nslots) return;
    BatchBlast* B = blasts + b;
    if (!B->active || B->ctl.done) return;
    const int lane = threadIdx.x;
    uint64_t k = B->ctl.red.slot[lane];
    B->ctl.red.slot[lane] = 0ull;
    k = wave_max_u64(k);
    if (lane != 0) return;
    const double dtn = B->ctl.C_min / key_speed(k);
    const unsigned long long it = B->ctl.it;
    const double t_new = B->ctl.t + dtn;
    B->ctl.dt = dtn;
    B->ctl.t = t_new;
    B->ctl.it = it + 1;
    B->cur ^= 1u;
    if (!isfinite(dtn)) {
        B->ctl.nonfinite = 1u;
        B->ctl.done = 1u;
    }
    if (!(t_new < B->ctl.max_t && it + 1 < B->ctl.max_it)) B->ctl.done = 1u;
    done_out[b] = static_cast<unsigned char>(B->ctl.done ? 1 + B->ctl.nonfinite : 0);
}

// Initial Dt of a blast whose IC was written on the device: min over the states present
// of their Dt, starting from 1e20 (hydro2d/main.cxx:287, 316-335).
__global__ void batch_dt0_kernel(BatchBlast* __restrict__ blasts, int b)
{
    BatchBlast* B = blasts + b;
    double m = 1e20;
    if (B->ic_flags & 1ull) m = (B->dt_in < m) ? B->dt_in : m;
    if (B->ic_flags & 2ull) m = (B->dt_out < m) ? B->dt_out : m;
    B->ctl.dt = m;
}

}  // namespace
}  // namespace hydro

using namespace hydro;

struct hydro_batch {
    int dim = 2;
    int precision = HYDRO_FP64;
    int device = 0;
    int cus = 256;
    size_t cap = 0;
    hipStream_t stream = nullptr;
    BatchBlast* d_blasts = nullptr;
    unsigned char* d_done = nullptr;
    std::vector<BatchBlast> h_blasts;  // host copies of the static fields (upload staging)
    std::vector<unsigned char> used;
    std::vector<unsigned char> h_done;
    // Segment table.
    BatchSeg* d_segs = nullptr;
    unsigned* d_off = nullptr;
    size_t seg_cap = 0, off_cap = 0;
    std::vector<BatchSeg> h_segs;
    std::vector<unsigned> h_off;
    bool dirty = true;
    // Tiling.
    int blocks_per_cu = 0;  // 0: default
    int tile_rows = 0;      // 0: balanced
    int group = 1;          // 2D: adjacent windows per worker (1, 2 or 4)
    unsigned grid = 0, nworkers = 0;
    bool tuned = false;
    // Graph of `graph_steps` steps.
    size_t graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    Consts<double> cs{};
};

namespace {

int max_blocks(const hydro_batch* b)
{
    if (b->dim == 3) return 3;
    return b->precision == HYDRO_FP32_FLUX ? kBatchWaves2<float> : kBatchWaves2<double>;
}

int default_blocks(const hydro_batch* b) { return b->dim == 2 ? max_blocks(b) : 2; }

void drop_graph(hydro_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// Columns of one blast: (count, rows each).
void blast_columns(const hydro_batch* b, const BatchBlast& B, int& ncol)
{
    if (b->dim == 2) {
        const int nwin = (B.nb_y + kOwned - 1) / kOwned;
        ncol = (nwin + b->group - 1) / b->group;
    } else {
        const int nzw = z_windows(B.nb_z);
        const int nyw = (B.nb_y + kTY - 1) / kTY;
        ncol = nzw * nyw;
    }
}

// Direct dispatch (2D, R-row tiles of one window): one wave per tile and a grid that
// covers the row-major tile table once, so the hardware dispatches tiles in table
// order and the rows in flight stay a narrow band, as in the plain step.  Persistent
// workers each walking a share of the table drift apart over a step; direct dispatch
// is 16% faster on the 256-blast mix and 30% on one 40000 x 20000 blast.
bool direct_mode(const hydro_batch* b) { return b->dim == 2 && b->tile_rows > 0 && b->group == 1; }

// Workgroups a direct step should have to fill the chip (8 per CU); tiles are halved
// down to 8 rows until there are that many, as the plain step's strips are.
constexpr long long kDirectFill = 2048;

int build_tables(hydro_batch* b)
{
    const int bpc = b->blocks_per_cu > 0 ? b->blocks_per_cu : default_blocks(b);
    b->grid = static_cast<unsigned>(bpc * b->cus);
    b->nworkers = b->dim == 2 ? b->grid * (kWavesPerBlock / b->group) : b->grid;
    const bool direct = direct_mode(b);
    const unsigned W = direct ? 1u : b->nworkers;
    std::vector<std::vector<BatchSeg>> per(W);
    // Concatenated columns in slot order.
    struct Col {
        int blast, col, rows;
    };
    std::vector<Col> cols;
    for (size_t s = 0; s < b->cap; ++s) {
        if (!b->used[s]) continue;
        int ncol = 0;
        blast_columns(b, b->h_blasts[s], ncol);
        for (int c = 0; c < ncol; ++c) cols.push_back({ static_cast<int>(s), c, b->h_blasts[s].nb_x });
    }
    if (b->tile_rows <= 0) {
        long long total = 0;
        for (const Col& c : cols) total += c.rows;
        // Worker w takes rows [w*total/W, (w+1)*total/W) of the concatenation.
        size_t ci = 0;
        long long base = 0;  // first concatenated row of cols[ci]
        for (unsigned w = 0; w < W && ci < cols.size(); ++w) {
            const long long lo = total * w / W, hi = total * (w + 1) / W;
            long long pos = lo;
            while (pos < hi && ci < cols.size()) {
                const long long cend = base + cols[ci].rows;
                const long long e = std::min(hi, cend);
                BatchSeg sg;
                sg.blast = cols[ci].blast;
                sg.col = cols[ci].col;
                sg.i0 = static_cast<int>(pos - base) + 1;
                sg.i1 = static_cast<int>(e - base);
                per[w].push_back(sg);
                pos = e;
                if (pos == cend) {
                    base = cend;
                    ++ci;
                }
            }
        }
    } else {
        // Tiles in row-major order within each blast (all columns of tile row 0, then tile
        // row 1, ...), dealt round-robin: the workers running at the same time hold
        // horizontally adjacent windows of the same rows, as the plain step's workgroups
        // do, so halo columns and DRAM pages are shared while they are in L2.  Column-major
        // order (each column's tiles in turn) ran 6% slower on the 256-blast mix.
        int R = b->tile_rows;
        if (direct) {
            auto tiles = [&](int r) {
                long long n = 0;
                for (const Col& c : cols) n += (c.rows + r - 1) / r;
                return n;
            };
            while (R > 8 && (tiles(R) + kWavesPerBlock - 1) / kWavesPerBlock < kDirectFill) R = std::max(8, R / 2);
        }
        size_t t = 0;
        for (size_t c0 = 0; c0 < cols.size();) {
            size_t c1 = c0;
            while (c1 < cols.size() && cols[c1].blast == cols[c0].blast) ++c1;
            for (int i0 = 1; i0 <= cols[c0].rows; i0 += R)
                for (size_t ci = c0; ci < c1; ++ci) {
                    BatchSeg sg;
                    sg.blast = cols[ci].blast;
                    sg.col = cols[ci].col;
                    sg.i0 = i0;
                    sg.i1 = std::min(i0 + R - 1, cols[ci].rows);
                    // Direct dispatch: odd tile rows march downward (i0 > i1 in the table).
                    if (direct && ((i0 - 1) / R) % 2 == 1) std::swap(sg.i0, sg.i1);
                    per[t % W].push_back(sg);
                    ++t;
                }
            c0 = c1;
        }
    }
    b->h_segs.clear();
    b->h_off.assign(W + 1, 0u);
    for (unsigned w = 0; w < W; ++w) {
        b->h_off[w] = static_cast<unsigned>(b->h_segs.size());
        b->h_segs.insert(b->h_segs.end(), per[w].begin(), per[w].end());
    }
    b->h_off[W] = static_cast<unsigned>(b->h_segs.size());
    if (direct) {
        const unsigned n = static_cast<unsigned>(b->h_segs.size());
        b->grid = std::max(1u, (n + kWavesPerBlock - 1) / kWavesPerBlock);
        b->nworkers = b->grid * kWavesPerBlock;
    }
    if (b->h_segs.empty()) b->h_segs.push_back(BatchSeg{ 0, 0, 1, 0 });  // keeps the buffer non-empty

    bool realloc = false;
    if (b->h_segs.size() > b->seg_cap) {
        if (b->d_segs) (void)hipFree(b->d_segs);
        b->d_segs = nullptr;
        const size_t n = b->h_segs.size() + b->h_segs.size() / 2 + 64;
        HYDRO_HIP_OR(hipMalloc(reinterpret_cast<void**>(&b->d_segs), n * sizeof(BatchSeg)), HYDRO_ERR_HIP);
        b->seg_cap = n;
        realloc = true;
    }
    if (b->h_off.size() > b->off_cap) {
        if (b->d_off) (void)hipFree(b->d_off);
        b->d_off = nullptr;
        HYDRO_HIP_OR(hipMalloc(reinterpret_cast<void**>(&b->d_off), b->h_off.size() * sizeof(unsigned)), HYDRO_ERR_HIP);
        b->off_cap = b->h_off.size();
        realloc = true;
    }
    HYDRO_HIP_OR(hipMemcpyAsync(b->d_segs, b->h_segs.data(), b->h_segs.size() * sizeof(BatchSeg),
                                hipMemcpyHostToDevice, b->stream),
                 HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipMemcpyAsync(b->d_off, b->h_off.data(), b->h_off.size() * sizeof(unsigned), hipMemcpyHostToDevice,
                                b->stream),
                 HYDRO_ERR_HIP);
    // The host staging vectors must outlive the copies.
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    if (realloc || direct) drop_graph(b);  // direct: the grid follows the table
    b->dirty = false;
    return 0;
}

int launch_step(hydro_batch* b)
{
    const dim3 grid(b->grid);
    if (b->dim == 2) {
        // Direct: blocks_per_cu caps the resident workgroups per CU (values below 3 act as
        // 3, the smallest cap the LDS reservation can express; 0 leaves it to registers).
        const int direct = direct_mode(b) ? 1 : 0;
        const int occ = direct && b->blocks_per_cu > 0 ? std::max(3, b->blocks_per_cu) : 0;
        if (b->precision == HYDRO_FP64) {
            constexpr auto k = batch2d_kernel<double, double, IeeeMath>;
            hipLaunchKernelGGL(k, grid, dim3(kBlock), occupancy_lds<k>(occ), b->stream, b->d_blasts, b->d_segs,
                               b->d_off, b->cs, b->nworkers, b->group, direct);
        } else {
            constexpr auto k = batch2d_kernel<double, float, FastMath>;
            hipLaunchKernelGGL(k, grid, dim3(kBlock), occupancy_lds<k>(occ), b->stream, b->d_blasts, b->d_segs,
                               b->d_off, b->cs, b->nworkers, b->group, direct);
        }
    } else {
        if (b->precision == HYDRO_FP64)
            hipLaunchKernelGGL((batch3d_kernel<double, double, IeeeMath>), grid, dim3(kBlock3), 0, b->stream,
                               b->d_blasts, b->d_segs, b->d_off, b->cs, b->nworkers);
        else
            hipLaunchKernelGGL((batch3d_kernel<double, float, FastMath>), grid, dim3(kBlock3), 0, b->stream,
                               b->d_blasts, b->d_segs, b->d_off, b->cs, b->nworkers);
    }
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(static_cast<unsigned>(b->cap)), dim3(kWave), 0, b->stream,
                       b->d_blasts, b->d_done, static_cast<unsigned>(b->cap));
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

int ensure_graph(hydro_batch* b)
{
    if (b->graph_valid) return 0;
    drop_graph(b);
    HYDRO_HIP_OR(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), HYDRO_ERR_HIP);
    int rc = 0;
    for (size_t n = 0; n < b->graph_steps && rc == 0; ++n) rc = launch_step(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("hydro_batch: graph capture failed: %s", hipGetErrorString(e));
        return HYDRO_ERR_HIP;
    }
    b->graph = g;
    HYDRO_HIP_OR(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), HYDRO_ERR_HIP);
    b->graph_valid = true;
    return 0;
}

int advance_impl(hydro_batch* b, size_t n_steps)
{
    if (b->dirty) {
        const int rc = build_tables(b);
        if (rc) return rc;
    }
    size_t n = n_steps;
    if (b->graph_steps > 0 && n >= b->graph_steps) {
        const int rc = ensure_graph(b);
        if (rc) return rc;
        for (; n >= b->graph_steps; n -= b->graph_steps)
            HYDRO_HIP_OR(hipGraphLaunch(b->exec, b->stream), HYDRO_ERR_HIP);
    }
    for (; n > 0; --n) {
        const int rc = launch_step(b);
        if (rc) return rc;
    }
    return 0;
}

bool check(hydro_batch* b, const char* fn)
{
    clear_error();
    if (!b) {
        set_error("%s: null batch", fn);
        return false;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != b->device) {
        set_error("%s: the batch belongs to device %d, the current device is %d", fn, b->device, dev);
        return false;
    }
    return true;
}

int add_common(hydro_batch* b, const size_t* nb, const double* D, double C_min, double max_t, size_t max_it,
               void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b, const char* fn)
{
    if (!nb || !D || !rhoE_a || !mom_a || !rhoE_b || !mom_b) {
        set_error("%s: null argument", fn);
        return HYDRO_ERR_ARG;
    }
    for (int d = 0; d < b->dim; ++d)
        if (nb[d] < 1 || nb[d] > (1u << 30)) {
            set_error("%s: nb[%d] = %zu out of [1, 2^30]", fn, d, nb[d]);
            return HYDRO_ERR_ARG;
        }
    if (!grid_fits(b->dim, nb)) {
        set_error("%s: grid of more than 2^40 cells", fn);
        return HYDRO_ERR_ARG;
    }
    size_t slot = 0;
    while (slot < b->cap && b->used[slot]) ++slot;
    if (slot == b->cap) {
        set_error("%s: batch is full (%zu blasts)", fn, b->cap);
        return HYDRO_ERR_NOMEM;
    }
    // The staging record of a reused slot may still be read by an earlier copy.
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    BatchBlast& B = b->h_blasts[slot];
    memset(&B, 0, sizeof(B));
    B.rhoE[0] = rhoE_a;
    B.rhoE[1] = rhoE_b;
    B.mom[0] = mom_a;
    B.mom[1] = mom_b;
    B.nb_x = static_cast<int>(nb[0]);
    B.nb_y = static_cast<int>(nb[1]);
    B.nb_z = b->dim == 3 ? static_cast<int>(nb[2]) : 1;
    B.active = 1;
    B.ys = static_cast<long long>(nb[1]) + 2;
    B.zs = b->dim == 3 ? static_cast<long long>(nb[2]) + 2 : 1;
    B.zy = b->dim == 3 ? B.ys * B.zs : B.ys;
    B.cur = 0;
    B.ctl.t = 0.0;
    B.ctl.max_t = max_t;
    B.ctl.it = 0;
    B.ctl.max_it = max_it;
    B.ctl.C_min = C_min;
    for (int d = 0; d < b->dim; ++d) B.ctl.D[d] = D[d];
    B.ctl.done = (0.0 < max_t && 0 < max_it) ? 0u : 1u;  // hydro2d/main.cxx:689
    b->used[slot] = 1;
    b->dirty = true;
    return static_cast<int>(slot);
}

int upload_record(hydro_batch* b, int slot)
{
    HYDRO_HIP_OR(hipMemcpyAsync(b->d_blasts + slot, &b->h_blasts[slot], sizeof(BatchBlast), hipMemcpyHostToDevice,
                                b->stream),
                 HYDRO_ERR_HIP);
    const unsigned char z = b->h_blasts[slot].ctl.done ? 1 : 0;
    b->h_done[slot] = z;
    HYDRO_HIP_OR(hipMemcpyAsync(b->d_done + slot, &b->h_done[slot], 1, hipMemcpyHostToDevice, b->stream), HYDRO_ERR_HIP);
    return 0;
}

// Autotuner cache: signature -> tiling.
struct TuneChoice {
    int bpc, rows, group;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune_cache;

std::string signature(const hydro_batch* b)
{
    // Dimension, precision, blast count and total cells (to 2 significant digits).
    long long cells = 0, n = 0;
    for (size_t s = 0; s < b->cap; ++s)
        if (b->used[s]) {
            const BatchBlast& B = b->h_blasts[s];
            cells += static_cast<long long>(B.nb_x) * B.nb_y * B.nb_z;
            ++n;
        }
    long long mag = 1;
    while (cells / mag >= 100) mag *= 10;
    char buf[96];
    snprintf(buf, sizeof buf, "d%d p%d n%lld c%lld", b->dim, b->precision, n, (cells / mag) * mag);
    return buf;
}

}  // namespace

extern "C" hydro_batch* hydro_batch_create(int dim, int precision, size_t capacity)
try {
    clear_error();
    if ((dim != 2 && dim != 3) || (precision != HYDRO_FP64 && precision != HYDRO_FP32_FLUX) || capacity < 1 ||
        capacity > (1u << 20)) {
        set_error("hydro_batch_create: dim must be 2 or 3, precision HYDRO_FP64 or HYDRO_FP32_FLUX, capacity in "
                  "[1, 2^20]");
        return nullptr;
    }
    hydro_batch* b = new (std::nothrow) hydro_batch();
    if (!b) {
        set_error("hydro_batch_create: out of host memory");
        return nullptr;
    }
    b->dim = dim;
    b->precision = precision;
    b->cap = capacity;
    try {
        b->h_blasts.assign(capacity, BatchBlast{});
        b->used.assign(capacity, 0);
        b->h_done.assign(capacity, 0);
    } catch (...) {
        delete b;  // nothing on the device yet
        throw;
    }
    const double gamma = 1.4, K = 1.1, gmo = gamma - 1.0;  // hydro2d/main.cxx:601-606
    b->cs.C_min = 0.0;
    b->cs.gamma = gamma;
    b->cs.gmo = gmo;
    b->cs.divgamma = 1.0 / gmo;
    b->cs.K = K;
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_blasts), capacity * sizeof(BatchBlast));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_done), capacity);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_blasts, 0, capacity * sizeof(BatchBlast), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_done, 0, capacity, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        set_error("hydro_batch_create: %s", hipGetErrorString(e));
        hydro_batch_destroy(b);
        return nullptr;
    }
    return b;
} catch (...) {
    hydro::set_exception_error(__func__);
    return nullptr;
}

extern "C" void hydro_batch_destroy(hydro_batch* b)
try {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->d_blasts) (void)hipFree(b->d_blasts);
    if (b->d_done) (void)hipFree(b->d_done);
    if (b->d_segs) (void)hipFree(b->d_segs);
    if (b->d_off) (void)hipFree(b->d_off);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
} catch (...) {
    hydro::set_exception_error(__func__);
}

extern "C" void* hydro_batch_stream(hydro_batch* b) { return b ? static_cast<void*>(b->stream) : nullptr; }

extern "C" int hydro_batch_add(hydro_batch* b, const size_t* nb, const double* D, double C_min, double dt0,
                               double max_t, size_t max_it, void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b)
try {
    if (!check(b, "hydro_batch_add")) return HYDRO_ERR_ARG;
    const int slot = add_common(b, nb, D, C_min, max_t, max_it, rhoE_a, mom_a, rhoE_b, mom_b, "hydro_batch_add");
    if (slot < 0) return slot;
    b->h_blasts[slot].ctl.dt = dt0;
    const int rc = upload_record(b, slot);
    if (rc) {
        b->used[slot] = 0;
        return rc;
    }
    return slot;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_add_blast(hydro_batch* b, const size_t* nb, double C, double radius_frac, double rho_in,
                                     double rho_out, double p_in, double p_out, double max_t, size_t max_it,
                                     void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b)
try {
    if (!check(b, "hydro_batch_add_blast")) return HYDRO_ERR_ARG;
    if (!nb) {
        set_error("hydro_batch_add_blast: null nb");
        return HYDRO_ERR_ARG;
    }
    if (!(radius_frac >= 0.0) || !(rho_in > 0.0) || !(rho_out > 0.0) || !(p_in > 0.0) || !(p_out > 0.0) ||
        !(C > 0.0)) {
        set_error("hydro_batch_add_blast: densities, pressures and C must be positive, radius_frac >= 0");
        return HYDRO_ERR_ARG;
    }
    // Spacings and C_min as hydro2d/main.cxx:482-484, 595-607: D = 1 / (NB + 1).
    double D[3] = { 0, 0, 0 };
    double ms = 0;
    for (int d = 0; d < b->dim; ++d) {
        D[d] = 1.0 / static_cast<double>(nb[d] + 1);
        ms = (d == 0 || D[d] < ms) ? D[d] : ms;
    }
    const double C_min = C * ms;
    const int slot = add_common(b, nb, D, C_min, max_t, max_it, rhoE_a, mom_a, rhoE_b, mom_b, "hydro_batch_add_blast");
    if (slot < 0) return slot;
    BatchBlast& B = b->h_blasts[slot];
    B.ic_flags = 0;
    int rc = 0;
    double dt_in = 0, dt_out = 0;
    // The IC's per-state Dt needs only host arithmetic; compute it first so that the
    // record is uploaded once, before the IC kernel or-s the flags into it.
    {
        rc = upload_record(b, slot);  // flags cleared on the device
        if (rc == 0)
            rc = ic_blast_async_f64(static_cast<double*>(rhoE_a), static_cast<double*>(mom_a), b->dim, nb, rho_in,
                                    rho_out, p_in, p_out, 1, radius_frac, C_min, 1.4, &b->d_blasts[slot].ic_flags,
                                    &dt_in, &dt_out, b->stream);
    }
    if (rc == 0) {
        B.dt_in = dt_in;
        B.dt_out = dt_out;
        // Only the two Dt fields change on the device record.
        hipError_t e = hipMemcpyAsync(&b->d_blasts[slot].dt_in, &B.dt_in, 2 * sizeof(double), hipMemcpyHostToDevice,
                                      b->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(batch_dt0_kernel, dim3(1), dim3(1), 0, b->stream, b->d_blasts, slot);
            e = hipGetLastError();
        }
        if (e != hipSuccess) {
            set_error("hydro_batch_add_blast: %s", hipGetErrorString(e));
            rc = HYDRO_ERR_HIP;
        }
    }
    if (rc) {
        b->used[slot] = 0;
        return rc;
    }
    return slot;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_remove(hydro_batch* b, int slot)
try {
    if (!check(b, "hydro_batch_remove")) return HYDRO_ERR_ARG;
    if (slot < 0 || static_cast<size_t>(slot) >= b->cap || !b->used[slot]) {
        set_error("hydro_batch_remove: slot %d is not in use", slot);
        return HYDRO_ERR_ARG;
    }
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    BatchBlast& B = b->h_blasts[slot];
    memset(&B, 0, sizeof(B));
    B.ctl.done = 1u;
    b->used[slot] = 0;
    b->dirty = true;
    const int rc = upload_record(b, slot);
    if (rc) return rc;
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_advance(hydro_batch* b, size_t n_steps)
try {
    if (!check(b, "hydro_batch_advance")) return HYDRO_ERR_ARG;
    return advance_impl(b, n_steps);
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_sync(hydro_batch* b)
try {
    if (!check(b, "hydro_batch_sync")) return HYDRO_ERR_ARG;
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_status(hydro_batch* b, int slot, size_t* it, double* t, double* dt, int* done,
                                  int* current)
try {
    if (!check(b, "hydro_batch_status")) return HYDRO_ERR_ARG;
    if (slot < 0 || static_cast<size_t>(slot) >= b->cap || !b->used[slot]) {
        set_error("hydro_batch_status: slot %d is not in use", slot);
        return HYDRO_ERR_ARG;
    }
    BatchBlast h;
    HYDRO_HIP_OR(hipMemcpyAsync(&h, b->d_blasts + slot, sizeof(h), hipMemcpyDeviceToHost, b->stream), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    if (it) *it = static_cast<size_t>(h.ctl.it);
    if (t) *t = h.ctl.t;
    if (dt) *dt = h.ctl.dt;
    if (done) *done = static_cast<int>(h.ctl.done);
    if (current) *current = static_cast<int>(h.cur);
    if (h.ctl.nonfinite) {
        set_error("hydro_batch_status: slot %d: Dt became non-finite at iteration %llu", slot, h.ctl.it);
        return HYDRO_ERR_NONFINITE;
    }
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_done_flags(hydro_batch* b, unsigned char* flags, size_t n)
try {
    if (!check(b, "hydro_batch_done_flags")) return HYDRO_ERR_ARG;
    if (!flags || n > b->cap) {
        set_error("hydro_batch_done_flags: null buffer or n > capacity");
        return HYDRO_ERR_ARG;
    }
    HYDRO_HIP_OR(hipMemcpyAsync(flags, b->d_done, n, hipMemcpyDeviceToHost, b->stream), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipStreamSynchronize(b->stream), HYDRO_ERR_HIP);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_set_tiling(hydro_batch* b, int blocks_per_cu, int tile_rows, int tile_windows)
try {
    if (!check(b, "hydro_batch_set_tiling")) return HYDRO_ERR_ARG;
    const bool width_ok = tile_windows == 0 || tile_windows == 1 || (b->dim == 2 && (tile_windows == 2 || tile_windows == 4));
    if (blocks_per_cu < 0 || blocks_per_cu > max_blocks(b) || tile_rows < 0 || tile_rows > (1 << 30) || !width_ok) {
        set_error("hydro_batch_set_tiling: blocks_per_cu in [0, %d] (0 = default), tile_rows >= 0 (0 = balanced), "
                  "tile_windows 0/1 (2D also 2 or 4)",
                  max_blocks(b));
        return HYDRO_ERR_ARG;
    }
    const int group = tile_windows > 0 ? tile_windows : 1;
    const bool was_direct = direct_mode(b);
    if (blocks_per_cu != b->blocks_per_cu || group != b->group) drop_graph(b);
    b->blocks_per_cu = blocks_per_cu;
    b->tile_rows = tile_rows;
    b->group = group;
    if (direct_mode(b) != was_direct) drop_graph(b);  // the grid and arguments differ
    b->dirty = true;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_get_tiling(hydro_batch* b, int* blocks_per_cu, int* tile_rows, int* tile_windows)
try {
    if (!check(b, "hydro_batch_get_tiling")) return HYDRO_ERR_ARG;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu > 0 ? b->blocks_per_cu : default_blocks(b);
    if (tile_rows) *tile_rows = b->tile_rows;
    if (tile_windows) *tile_windows = b->group;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_set_graph(hydro_batch* b, size_t steps_per_graph)
try {
    if (!check(b, "hydro_batch_set_graph")) return HYDRO_ERR_ARG;
    if (steps_per_graph > 4096) {
        set_error("hydro_batch_set_graph: at most 4096 steps per graph");
        return HYDRO_ERR_ARG;
    }
    if (steps_per_graph != b->graph_steps) drop_graph(b);
    b->graph_steps = steps_per_graph;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_batch_autotune(hydro_batch* b, size_t steps_per_trial, int use_cache, double* best_ms_per_step)
try {
    if (!check(b, "hydro_batch_autotune")) return HYDRO_ERR_ARG;
    if (steps_per_trial < 1) {
        set_error("hydro_batch_autotune: steps_per_trial must be >= 1");
        return HYDRO_ERR_ARG;
    }
    const std::string sig = signature(b);
    if (use_cache) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        const auto it = g_tune_cache.find(sig);
        if (it != g_tune_cache.end()) {
            if (best_ms_per_step) *best_ms_per_step = -1.0;
            return hydro_batch_set_tiling(b, it->second.bpc, it->second.rows, it->second.group);
        }
    }
    std::vector<TuneChoice> cands;
    if (b->dim == 2) {
        // Direct dispatch: workgroups-per-CU cap x tile rows; the balanced persistent
        // partition: grid occupancy.  (Multi-window tiles were never the fastest.)
        for (int occ = 3; occ <= max_blocks(b); ++occ)
            for (int r : { 16, 32, 64 }) cands.push_back({ occ, r, 1 });
        for (int bpc = 2; bpc <= max_blocks(b); ++bpc) cands.push_back({ bpc, 0, 1 });
    } else {
        for (int bpc = 1; bpc <= max_blocks(b); ++bpc)
            for (int r : { 0, 4, 16 }) cands.push_back({ bpc, r, 1 });
    }
    // Trials use direct launches; the graph setting and the events are restored and
    // released on every way out, errors and exceptions included.
    struct TrialScope {
        hydro_batch* b;
        size_t graph_steps;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~TrialScope()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            b->graph_steps = graph_steps;
        }
    } scope{ b, b->graph_steps };
    b->graph_steps = 0;
    HYDRO_HIP_OR(hipEventCreate(&scope.e0), HYDRO_ERR_HIP);
    HYDRO_HIP_OR(hipEventCreate(&scope.e1), HYDRO_ERR_HIP);
    hipEvent_t e0 = scope.e0, e1 = scope.e1;
    auto trial = [&](const TuneChoice& c, double& per) -> int {
        int rc = hydro_batch_set_tiling(b, c.bpc, c.rows, c.group);
        if (rc == 0 && b->dirty) rc = build_tables(b);
        if (rc) return rc;
        hipError_t e = hipEventRecord(e0, b->stream);
        if (e == hipSuccess) rc = advance_impl(b, steps_per_trial);
        if (e == hipSuccess && rc == 0) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess && rc == 0) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess && rc == 0) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            set_error("hydro_batch_autotune: %s", hipGetErrorString(e));
            return HYDRO_ERR_HIP;
        }
        per = ms / static_cast<double>(steps_per_trial);
        return rc;
    };
    // One timed trial per candidate, then the three fastest again; keep the fastest
    // second timing (a single short trial is noisy).
    std::vector<std::pair<double, size_t>> first;
    int rc = 0;
    for (size_t i = 0; i < cands.size() && rc == 0; ++i) {
        double per = 0;
        rc = trial(cands[i], per);
        first.push_back({ per, i });
    }
    double best = 1e300;
    TuneChoice best_c = cands.empty() ? TuneChoice{ 0, 0, 1 } : cands[0];
    if (rc == 0) {
        std::sort(first.begin(), first.end());
        for (size_t k = 0; k < first.size() && k < 3 && rc == 0; ++k) {
            double per = 0;
            rc = trial(cands[first[k].second], per);
            if (per < best) {
                best = per;
                best_c = cands[first[k].second];
            }
        }
    }
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[sig] = best_c;
    }
    if (best_ms_per_step) *best_ms_per_step = best;
    return hydro_batch_set_tiling(b, best_c.bpc, best_c.rows, best_c.group);
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}
