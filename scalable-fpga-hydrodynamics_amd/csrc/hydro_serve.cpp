// Served form: a request loop over the batched engines with continuous batching.
//
// Requests (one blast solve each: dimension, size, precision, blast parameters, stop
// conditions) are queued on arrival and admitted in arrival order as soon as their
// device footprint (two FP64 buffer pairs, ghosts included) fits in the server's HBM
// arena and their engine has a free slot.  One batch engine per (dimension, precision)
// advances all of its running requests together on its own stream, so requests of every
// kind overlap on the device; a request joins its engine between that engine's ticks and
// leaves it when its own loop ends (max_it or max_t, hydro2d/main.cxx:689), freeing
// memory for the queue.  State stays on the device from the initial condition (written
// there) to the end; a result is copied out only if the client asks for it.
//
// Engines tick independently: each has an event marking the end of its last tick, and
// hydro_server_tick serves every engine that is idle (retire its finished requests,
// admit queued ones, enqueue its next steps_per_tick steps), waiting only when every
// engine with work is still busy.  A small engine therefore never waits for a large
// one.  Latency = completion observed after the request's last tick - arrival, on the
// server's steady clock.
//
// Size classes (hydro_server_set_policy): a request whose estimated solve time alone
// (its steps times a per-step cost of the plain resident loop on MI355X, launch floor
// plus cells) reaches `large_seconds` is large.  Large requests run one at a time, in
// arrival order, each as a plain device-resident solve (hydro_run) on a stream of their
// own; smaller requests go to the batched engines at once, beside it.  A large solve's
// steps reach its stream in chunks of about kLargeChunkSeconds of work, at most two in
// flight: HIP maps streams onto a few hardware queues (4 by default), and a queue shared
// with a batched engine's stream then interleaves the two instead of holding the
// engine's ticks behind the whole solve (enqueued at once, a second server in one
// process ran at p99 0.84 s instead of 0.48 s on the served mix, and every server did
// with 16 hardware queues).  Sharing the device among everything admitted (the default, one
// class) stretches the largest solves by the load, and a plain FIFO makes small solves
// wait behind large ones; the size classes keep the large solves close to their FIFO
// latency while the small ones overtake them.  Under a backlog (hydro_server_set_backlog:
// more than `large_backlog` large requests waiting), the oldest waiting large requests
// join their batched engines as small ones do: one plain solve at a time is the slowest
// way through a long queue of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <new>
#include <string.h>
#include <thread>
#include <vector>

#include "../../include/hydro_hip.h"
#include "hydro_kernels.hpp"
#include "hydro_runtime.hpp"

using namespace hydro;

namespace {

constexpr size_t kAlign = 256;

// First-fit allocator over [0, size) with coalescing free list (offset -> length).
class Arena {
public:
    void reset(size_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t& off)
    {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0;
};

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kReleased = 3, kFailed = 4 };

struct Req {
    int dim = 2;
    size_t nb[3] = { 1, 1, 1 };
    int precision = HYDRO_FP64;
    double C = 0.5, rf = 0.1, rho_in = 1.0, rho_out = 1.2, p_in = 10.0, p_out = 0.1, max_t = 1e30;
    size_t max_it = 0;
    bool keep = false;
    int state = kQueued;
    double t_arrival = 0, t_admit = -1, t_done = -1;
    size_t off = 0, bytes = 0, buf_bytes[2] = { 0, 0 };  // rhoE / momentum bytes per buffer
    int engine = -1, slot = -1, current = 0, rc = 0;
    size_t it = 0;
    double t_sim = 0;
    double interior = 0;  // interior cells
    bool large = false;   // in the large size class (runs alone among large requests)
};

// Estimated time of a request's solve alone on the plain resident loop: per step a
// launch floor plus a cost per interior cell, fitted to plain solves of the served mix on
// MI355X (2D: 9 us + 13 ps per cell, i.e. about 4.9 TB/s at 64 B per cell; 3D: 11 us +
// 24 ps per cell, about 3.4 TB/s at 80 B).  Only ranks requests into the size classes.
double solo_seconds(int dim, double interior, size_t steps)
{
    const double floor_s = dim == 3 ? 11e-6 : 9e-6;
    const double per_cell = dim == 3 ? 24e-12 : 13e-12;
    return static_cast<double>(steps) * (floor_s + per_cell * interior);
}

size_t cells_of(const Req& r)
{
    size_t c = 1;
    for (int d = 0; d < r.dim; ++d) c *= r.nb[d] + 2;
    return c;
}

// Batched engines, one per (dimension, precision); large requests use none.
constexpr int kEngines = 4;
int engine_of(const Req& r) { return (r.dim == 3 ? 2 : 0) + (r.precision == HYDRO_FP32_FLUX ? 1 : 0); }

}  // namespace

struct hydro_server {
    int device = 0;
    char* base = nullptr;
    Arena arena;
    size_t capacity = 0;  // running requests per engine
    size_t steps_per_tick = 50;
    hydro_batch* eng[kEngines] = {};
    hipEvent_t ev[kEngines] = {};  // end of the engine's last tick
    bool busy[kEngines] = {};
    std::vector<int> slot_req[kEngines];  // slot -> request id (-1 free)
    size_t running[kEngines] = {};
    std::vector<Req> reqs;
    std::deque<int> queue;
    std::chrono::steady_clock::time_point t0;
    size_t ticks = 0;
    double cell_updates = 0;  // of retired requests
    double max_active_cells = 0;  // admission cap on interior cells in flight per engine (0 = none)
    double active_cells[kEngines] = {};
    // Size classes: threshold (0: one class) and the one large request running, if any.
    double large_seconds = 0;
    size_t large_backlog = 0;  // 0: large requests never join the batched engines
    int large_id = -1;
    hydro_run* large_run = nullptr;
    hipStream_t large_stream = nullptr;
    hipEvent_t large_ev[2] = {};  // end of chunk k of the large solve in large_ev[k % 2]
    size_t large_steps = 0;       // steps of the running large request
    size_t large_enq = 0;         // of them enqueued so far
    size_t large_chunk = 1;       // steps per chunk
    size_t large_chunks = 0;      // chunks enqueued so far

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

bool check(hydro_server* s, const char* fn)
{
    clear_error();
    if (!s) {
        set_error("%s: null server", fn);
        return false;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != s->device) {
        set_error("%s: the server belongs to device %d, the current device is %d", fn, s->device, dev);
        return false;
    }
    return true;
}

int ensure_engine(hydro_server* s, int e)
{
    if (s->eng[e]) return 0;
    const int dim = e >= 2 ? 3 : 2;
    const int prec = (e & 1) ? HYDRO_FP32_FLUX : HYDRO_FP64;
    // Built in locals and published only once complete, so that a failure leaves no
    // half-configured engine for the next call to take as ready.
    hydro_batch* b = hydro_batch_create(dim, prec, s->capacity);
    if (!b) return HYDRO_ERR_HIP;
    // Continuous batching replays a few steps at a time from a graph.  The balanced
    // tiling (equal row ranges per worker of a persistent grid) stays efficient whatever
    // the mix, down to a single request; measured on the served mix, it beats direct
    // dispatch of R-row tiles (p50 57 vs 71-83 ms).  Two workgroups per CU per engine:
    // with four engines sharing the device, three per CU for the 2D engines gave a p50
    // of 58 ms against 52 ms at the same throughput.
    int rc = hydro_batch_set_graph(b, 10);
    if (rc == 0) rc = hydro_batch_set_tiling(b, 2, 0, 1);
    hipEvent_t ev = nullptr;
    if (rc == 0 && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        set_error("hydro server: cannot create an event");
        rc = HYDRO_ERR_HIP;
    }
    if (rc == 0) {
        try {
            s->slot_req[e].assign(s->capacity, -1);
        } catch (...) {
            set_error("hydro server: out of host memory");
            rc = HYDRO_ERR_NOMEM;
        }
    }
    if (rc != 0) {
        if (ev) (void)hipEventDestroy(ev);
        hydro_batch_destroy(b);
        return rc;
    }
    s->eng[e] = b;
    s->ev[e] = ev;
    return 0;
}

// The four arrays of a request in its arena block: rhoE and momentum of buffer a, then
// of buffer b.
void buffers_of(hydro_server* s, const Req& r, size_t off, void* p[4])
{
    char* b = s->base + off;
    p[0] = b;
    p[1] = b + r.buf_bytes[0];
    p[2] = b + r.buf_bytes[0] + r.buf_bytes[1];
    p[3] = b + 2 * r.buf_bytes[0] + r.buf_bytes[1];
}

// Start a large request as a plain resident solve on the large stream: the server's one
// resident run is restarted for it (the blast IC of hydro_batch_add_blast, ic_mode 1,
// D = 1 / (nb + 1), C_min = C * min(D), written on the device) and its first chunk of
// steps is enqueued (feed_large).  Nothing is allocated and nothing waits, except when a
// request needs a longer (t, Dt) record than the run has (created with at least
// kLargeRecord iterations), which reallocates it once.  r.off is set.
constexpr size_t kLargeRecord = 4096;
constexpr double kLargeChunkSeconds = 2e-3;

// Enqueue the running large solve's next chunk of steps and record its event.
int feed_large(hydro_server* s)
{
    const Req& r = s->reqs[s->large_id];
    void* p[4];
    buffers_of(s, r, r.off, p);
    const size_t n = std::min(s->large_chunk, s->large_steps - s->large_enq);
    int rc = hydro_run_advance(s->large_run, p[0], p[1], p[2], p[3], n, s->large_stream);
    if (rc == 0 && hipEventRecord(s->large_ev[s->large_chunks % 2], s->large_stream) != hipSuccess) {
        set_error("hydro server: cannot record the large solve's event");
        rc = HYDRO_ERR_HIP;
    }
    if (rc == 0) {
        s->large_enq += n;
        ++s->large_chunks;
    }
    return rc;
}

// What the running large solve needs: 1 its next chunk (steps remain and chunk k - 2,
// whose event slot chunk k takes, has completed), 2 retiring (every step has run), 0
// nothing yet; or a HIP error code.
int large_state(hydro_server* s)
{
    const size_t k = s->large_chunks;  // at least 1 once started
    const bool more = s->large_enq < s->large_steps;
    if (more && k < 2) return 1;
    const hipError_t q = hipEventQuery(s->large_ev[more ? k % 2 : (k - 1) % 2]);
    if (q == hipSuccess) return more ? 1 : 2;
    if (q != hipErrorNotReady) {
        set_error("hydro server: large solve: %s", hipGetErrorString(q));
        return HYDRO_ERR_HIP;
    }
    return 0;
}

int start_large(hydro_server* s, int id, size_t off)
{
    Req& r = s->reqs[id];
    void* p[4];
    buffers_of(s, r, off, p);
    if (!s->large_run || hydro::run_capacity(s->large_run) < r.max_it) {
        if (s->large_run) {
            (void)hipStreamSynchronize(s->large_stream);
            hydro_run_destroy(s->large_run);
            s->large_run = nullptr;
        }
        const size_t nb1[3] = { 1, 1, 1 };
        const double D1[3] = { 0.5, 0.5, 0.5 };
        s->large_run = hydro_run_create(2, nb1, D1, 0.1, 1e-3, 1.0, r.max_it > kLargeRecord ? r.max_it : kLargeRecord,
                                        HYDRO_FP64, nullptr);
        if (!s->large_run) return HYDRO_ERR_HIP;
    }
    int rc = hydro::run_restart_blast(s->large_run, r.dim, r.nb, r.precision, r.C, r.rf, r.rho_in, r.rho_out, r.p_in,
                                      r.p_out, r.max_t, r.max_it, p[0], p[1], s->large_stream);
    if (rc == 0) {
        r.off = off;
        s->large_id = id;
        s->large_steps = r.max_it;
        s->large_enq = 0;
        s->large_chunks = 0;
        // Steps per chunk from the solve-time estimate of the size classes.
        const double per_step = solo_seconds(r.dim, r.interior, 1);
        const double c = std::ceil(kLargeChunkSeconds / per_step);
        s->large_chunk = c < 1.0 ? 1 : (c >= static_cast<double>(r.max_it) ? r.max_it : static_cast<size_t>(c));
        rc = feed_large(s);
    }
    if (rc) {
        (void)hipStreamSynchronize(s->large_stream);
        s->large_id = -1;
        return rc;
    }
    return 0;
}

size_t queued_large(const hydro_server* s)
{
    size_t n = 0;
    for (const int id : s->queue) n += s->reqs[id].large ? 1 : 0;
    return n;
}

// A waiting large request may join its batched engine: more than large_backlog wait.
bool over_backlog(const hydro_server* s, size_t waiting) { return s->large_backlog > 0 && waiting > s->large_backlog; }

// Admit queued requests in arrival order.  A large request starts on the plain resident
// run when no large request runs; otherwise it waits (later large requests wait behind
// it) unless the backlog lets it join its engine.  A request whose engine is busy
// (mid-tick), full or at its active-cell cap waits without holding back requests of other
// engines.  Admission stops at the first request whose memory does not fit, so that later
// requests never take the memory an earlier one is waiting for.
int admit(hydro_server* s)
{
    size_t waiting = queued_large(s);
    for (auto q = s->queue.begin(); q != s->queue.end();) {
        const int id = *q;
        Req& r = s->reqs[id];
        if (r.large && s->large_id < 0) {
            size_t off = 0;
            if (!s->arena.alloc(r.bytes, off)) break;
            const int rc = start_large(s, id, off);
            if (rc) {
                s->arena.release(off, r.bytes);
                return rc;
            }
            r.off = off;
            r.state = kRunning;
            r.t_admit = s->now();
            q = s->queue.erase(q);
            --waiting;
            continue;
        }
        if (r.large && !over_backlog(s, waiting)) {
            // One large request at a time, in arrival order (every later large request is
            // held back too); small ones pass.
            ++q;
            continue;
        }
        const int e = engine_of(r);
        int rc = ensure_engine(s, e);
        if (rc) return rc;
        if (s->busy[e] || s->running[e] >= s->capacity) {
            ++q;
            continue;
        }
        // The cap holds back this engine only: later requests of other engines may pass.
        if (s->max_active_cells > 0 && s->running[e] > 0 && s->active_cells[e] + r.interior > s->max_active_cells) {
            ++q;
            continue;
        }
        size_t off = 0;
        if (!s->arena.alloc(r.bytes, off)) break;
        void* p[4];
        buffers_of(s, r, off, p);
        const int slot = hydro_batch_add_blast(s->eng[e], r.nb, r.C, r.rf, r.rho_in, r.rho_out, r.p_in, r.p_out,
                                               r.max_t, r.max_it, p[0], p[1], p[2], p[3]);
        if (slot < 0) {
            s->arena.release(off, r.bytes);
            return slot;
        }
        r.off = off;
        r.engine = e;
        r.slot = slot;
        r.state = kRunning;
        r.t_admit = s->now();
        s->slot_req[e][slot] = id;
        ++s->running[e];
        s->active_cells[e] += r.interior;
        q = s->queue.erase(q);
        if (r.large) --waiting;
    }
    return 0;
}

void release_memory(hydro_server* s, Req& r)
{
    s->arena.release(r.off, r.bytes);
    r.state = r.state == kFailed ? kFailed : kReleased;
}

int retire(hydro_server* s, int e, size_t* n_done)
{
    std::vector<unsigned char> flags(s->capacity);
    {
        if (!s->eng[e] || s->running[e] == 0) return 0;
        int rc = hydro_batch_done_flags(s->eng[e], flags.data(), s->capacity);
        if (rc) return rc;
        const double t = s->now();
        for (size_t slot = 0; slot < s->capacity; ++slot) {
            const int id = s->slot_req[e][slot];
            if (id < 0 || !flags[slot]) continue;
            Req& r = s->reqs[id];
            int done = 0;
            r.rc = hydro_batch_status(s->eng[e], static_cast<int>(slot), &r.it, &r.t_sim, nullptr, &done, &r.current);
            r.t_done = t;
            r.state = r.rc == 0 ? kDone : kFailed;
            s->cell_updates += static_cast<double>(r.it) * r.interior;
            rc = hydro_batch_remove(s->eng[e], static_cast<int>(slot));
            if (rc) return rc;
            s->slot_req[e][slot] = -1;
            --s->running[e];
            s->active_cells[e] -= r.interior;
            if (!r.keep || r.state == kFailed) release_memory(s, r);
            if (n_done) ++*n_done;
        }
    }
    return 0;
}

// Retire the large request once its event has completed (the caller checked).
int retire_large(hydro_server* s, size_t* n_done)
{
    Req& r = s->reqs[s->large_id];
    int done = 0;
    r.rc = hydro_run_status(s->large_run, &r.it, &r.t_sim, nullptr, &done, &r.current, s->large_stream);
    r.t_done = s->now();
    r.state = r.rc == 0 ? kDone : kFailed;
    s->cell_updates += static_cast<double>(r.it) * r.interior;
    s->large_id = -1;
    if (!r.keep || r.state == kFailed) release_memory(s, r);
    if (n_done) ++*n_done;
    return 0;
}

}  // namespace

extern "C" hydro_server* hydro_server_create(size_t hbm_bytes, size_t capacity, size_t steps_per_tick)
try {
    clear_error();
    if (hbm_bytes < kAlign || capacity < 1 || capacity > (1u << 16) || steps_per_tick < 1) {
        set_error("hydro_server_create: hbm_bytes >= 256, capacity in [1, 65536], steps_per_tick >= 1");
        return nullptr;
    }
    hydro_server* s = new (std::nothrow) hydro_server();
    if (!s) {
        set_error("hydro_server_create: out of host memory");
        return nullptr;
    }
    s->capacity = capacity;
    s->steps_per_tick = steps_per_tick;
    s->t0 = std::chrono::steady_clock::now();
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->base), hbm_bytes);
    // The large class's stream (default priority) and completion event.
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->large_stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&s->large_ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("hydro_server_create: %s (%zu bytes)", hipGetErrorString(e), hbm_bytes);
        hydro_server_destroy(s);
        return nullptr;
    }
    try {
        s->arena.reset(hbm_bytes);
    } catch (...) {
        hydro_server_destroy(s);  // frees the arena's device memory and the server
        throw;
    }
    return s;
} catch (...) {
    hydro::set_exception_error(__func__);
    return nullptr;
}

extern "C" void hydro_server_destroy(hydro_server* s)
try {
    if (!s) return;
    for (int e = 0; e < kEngines; ++e) {
        if (s->eng[e]) hydro_batch_destroy(s->eng[e]);
        if (s->ev[e]) (void)hipEventDestroy(s->ev[e]);
    }
    if (s->large_stream) (void)hipStreamSynchronize(s->large_stream);
    if (s->large_run) hydro_run_destroy(s->large_run);
    for (hipEvent_t ev : s->large_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (s->large_stream) (void)hipStreamDestroy(s->large_stream);
    if (s->base) (void)hipFree(s->base);
    delete s;
} catch (...) {
    hydro::set_exception_error(__func__);
}

extern "C" double hydro_server_now(hydro_server* s) { return s ? s->now() : 0.0; }

extern "C" int hydro_server_submit(hydro_server* s, int dim, const size_t* nb, int precision, double C,
                                   double radius_frac, double rho_in, double rho_out, double p_in, double p_out,
                                   double max_t, size_t max_it, int keep_result, double arrival)
try {
    if (!check(s, "hydro_server_submit")) return HYDRO_ERR_ARG;
    if ((dim != 2 && dim != 3) || !nb || (precision != HYDRO_FP64 && precision != HYDRO_FP32_FLUX)) {
        set_error("hydro_server_submit: dim 2 or 3, nb non-null, precision HYDRO_FP64 or HYDRO_FP32_FLUX");
        return HYDRO_ERR_ARG;
    }
    Req r;
    r.dim = dim;
    for (int d = 0; d < dim; ++d) {
        if (nb[d] < 1 || nb[d] > (1u << 30)) {
            set_error("hydro_server_submit: nb[%d] = %zu out of [1, 2^30]", d, nb[d]);
            return HYDRO_ERR_ARG;
        }
        r.nb[d] = nb[d];
    }
    r.precision = precision;
    r.C = C;
    r.rf = radius_frac;
    r.rho_in = rho_in;
    r.rho_out = rho_out;
    r.p_in = p_in;
    r.p_out = p_out;
    r.max_t = max_t;
    r.max_it = max_it;
    r.keep = keep_result != 0;
    r.t_arrival = arrival >= 0 ? arrival : s->now();
    // The request's footprint, computed with overflow checks: a product that wrapped
    // could pass the arena check below with a short buffer that the kernels overrun.
    // Each of the four arrays starts on a 256-byte boundary of the arena.
    size_t cells = 1;
    bool wrapped = false;
    for (int d = 0; d < dim; ++d) wrapped |= __builtin_mul_overflow(cells, r.nb[d] + 2, &cells);
    size_t b0 = 0, b1 = 0;
    wrapped |= __builtin_mul_overflow(cells, size_t(2 * 8), &b0);
    wrapped |= __builtin_mul_overflow(cells, static_cast<size_t>(dim) * 8, &b1);
    if (wrapped || b0 > SIZE_MAX / 8 || b1 > SIZE_MAX / 8) {
        set_error("hydro_server_submit: request of %zu x %zu x %zu cells is too large", r.nb[0], r.nb[1],
                  dim == 3 ? r.nb[2] : size_t(1));
        return HYDRO_ERR_NOMEM;
    }
    r.interior = 1;
    for (int d = 0; d < dim; ++d) r.interior *= static_cast<double>(r.nb[d]);
    r.large = s->large_seconds > 0 && solo_seconds(dim, r.interior, max_it) >= s->large_seconds;
    r.buf_bytes[0] = (b0 + kAlign - 1) / kAlign * kAlign;
    r.buf_bytes[1] = (b1 + kAlign - 1) / kAlign * kAlign;
    r.bytes = 2 * (r.buf_bytes[0] + r.buf_bytes[1]);
    if (r.bytes > s->arena.size()) {
        set_error("hydro_server_submit: request needs %zu bytes, more than the server's %zu", r.bytes,
                  s->arena.size());
        return HYDRO_ERR_NOMEM;
    }
    // Queue first: if the request list cannot grow, the queue entry is taken back, so a
    // failed submit leaves neither a queued id without a request nor the reverse.
    const int id = static_cast<int>(s->reqs.size());
    s->queue.push_back(id);
    try {
        s->reqs.push_back(r);
    } catch (...) {
        s->queue.pop_back();
        throw;
    }
    return id;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_tick(hydro_server* s, size_t* n_done)
try {
    if (!check(s, "hydro_server_tick")) return HYDRO_ERR_ARG;
    if (n_done) *n_done = 0;
    // Wait until an engine that has work (running requests, or queued ones it could
    // admit) is idle.
    for (;;) {
        bool any_busy = false, any_ready = false;
        for (int e = 0; e < kEngines; ++e) {
            if (s->busy[e]) {
                const hipError_t q = hipEventQuery(s->ev[e]);
                if (q == hipSuccess) s->busy[e] = false;
                else if (q != hipErrorNotReady) {
                    set_error("hydro_server_tick: %s", hipGetErrorString(q));
                    return HYDRO_ERR_HIP;
                }
            }
            any_busy = any_busy || s->busy[e];
            any_ready = any_ready || (!s->busy[e] && s->running[e] > 0);
        }
        if (s->large_id >= 0) {
            const int ls = large_state(s);  // 1: takes its next chunk, 2: retires
            if (ls < 0) return ls;
            any_busy = any_busy || ls != 2;
            any_ready = any_ready || ls != 0;
        }
        const bool backlog = over_backlog(s, queued_large(s));
        for (const int id : s->queue) {
            const Req& r = s->reqs[id];
            const bool engine_free = !s->busy[engine_of(r)];
            any_ready = any_ready || (r.large ? s->large_id < 0 || (backlog && engine_free) : engine_free);
        }
        if (any_ready || !any_busy) break;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    int rc = 0;
    if (s->large_id >= 0) {
        int ls = large_state(s);
        while (ls == 1) {  // up to two chunks in flight
            rc = feed_large(s);
            ls = rc ? rc : large_state(s);
        }
        if (ls == 2) rc = retire_large(s, n_done);
        else if (ls < 0) rc = ls;
    }
    for (int e = 0; e < kEngines && rc == 0; ++e)
        if (!s->busy[e]) rc = retire(s, e, n_done);
    if (rc == 0) rc = admit(s);
    for (int e = 0; e < kEngines && rc == 0; ++e)
        if (s->eng[e] && !s->busy[e] && s->running[e] > 0) {
            rc = hydro_batch_advance(s->eng[e], s->steps_per_tick);
            if (rc == 0) {
                const hipError_t r = hipEventRecord(s->ev[e], static_cast<hipStream_t>(hydro_batch_stream(s->eng[e])));
                if (r != hipSuccess) {
                    set_error("hydro_server_tick: %s", hipGetErrorString(r));
                    rc = HYDRO_ERR_HIP;
                }
                s->busy[e] = true;
            }
        }
    ++s->ticks;
    return rc;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_pending(hydro_server* s, size_t* queued, size_t* running, size_t* bytes_in_use)
try {
    if (!check(s, "hydro_server_pending")) return HYDRO_ERR_ARG;
    if (queued) *queued = s->queue.size();
    if (running) {
        *running = s->large_id >= 0 ? 1 : 0;
        for (int e = 0; e < kEngines; ++e) *running += s->running[e];
    }
    if (bytes_in_use) *bytes_in_use = s->arena.used();
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_request(hydro_server* s, int id, int* state, double* t_arrival, double* t_admit,
                                    double* t_done, size_t* it, double* t_sim)
try {
    if (!check(s, "hydro_server_request")) return HYDRO_ERR_ARG;
    if (id < 0 || static_cast<size_t>(id) >= s->reqs.size()) {
        set_error("hydro_server_request: unknown request %d", id);
        return HYDRO_ERR_ARG;
    }
    const Req& r = s->reqs[id];
    if (state) *state = r.state;
    if (t_arrival) *t_arrival = r.t_arrival;
    if (t_admit) *t_admit = r.t_admit;
    if (t_done) *t_done = r.t_done;
    if (it) *it = r.it;
    if (t_sim) *t_sim = r.t_sim;
    return r.rc;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_result(hydro_server* s, int id, void* dst_rhoE, void* dst_mom)
try {
    if (!check(s, "hydro_server_result")) return HYDRO_ERR_ARG;
    if (id < 0 || static_cast<size_t>(id) >= s->reqs.size() || s->reqs[id].state != kDone) {
        set_error("hydro_server_result: request %d has no kept result", id);
        return HYDRO_ERR_ARG;
    }
    const Req& r = s->reqs[id];
    const char* p = s->base + r.off;
    const char* rhoE = p + (r.current ? r.buf_bytes[0] + r.buf_bytes[1] : 0);
    const char* mom = rhoE + r.buf_bytes[0];
    const size_t cells = cells_of(r);
    if (dst_rhoE) HYDRO_HIP_OR(hipMemcpy(dst_rhoE, rhoE, cells * 2 * 8, hipMemcpyDefault), HYDRO_ERR_HIP);
    if (dst_mom)
        HYDRO_HIP_OR(hipMemcpy(dst_mom, mom, cells * static_cast<size_t>(r.dim) * 8, hipMemcpyDefault), HYDRO_ERR_HIP);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_release(hydro_server* s, int id)
try {
    if (!check(s, "hydro_server_release")) return HYDRO_ERR_ARG;
    if (id < 0 || static_cast<size_t>(id) >= s->reqs.size() || s->reqs[id].state != kDone) {
        set_error("hydro_server_release: request %d has no kept result", id);
        return HYDRO_ERR_ARG;
    }
    release_memory(s, s->reqs[id]);
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_set_limits(hydro_server* s, double max_active_cells, size_t steps_per_tick)
try {
    if (!check(s, "hydro_server_set_limits")) return HYDRO_ERR_ARG;
    if (!(max_active_cells >= 0) || steps_per_tick < 1) {
        set_error("hydro_server_set_limits: max_active_cells >= 0, steps_per_tick >= 1");
        return HYDRO_ERR_ARG;
    }
    s->max_active_cells = max_active_cells;
    s->steps_per_tick = steps_per_tick;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_set_policy(hydro_server* s, double large_seconds)
try {
    if (!check(s, "hydro_server_set_policy")) return HYDRO_ERR_ARG;
    if (!(large_seconds >= 0)) {
        set_error("hydro_server_set_policy: large_seconds >= 0");
        return HYDRO_ERR_ARG;
    }
    s->large_seconds = large_seconds;  // classifies requests submitted from now on
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_set_backlog(hydro_server* s, size_t large_backlog)
try {
    if (!check(s, "hydro_server_set_backlog")) return HYDRO_ERR_ARG;
    s->large_backlog = large_backlog;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}

extern "C" int hydro_server_stats(hydro_server* s, size_t* ticks, double* cell_updates)
try {
    if (!check(s, "hydro_server_stats")) return HYDRO_ERR_ARG;
    if (ticks) *ticks = s->ticks;
    if (cell_updates) *cell_updates = s->cell_updates;
    return 0;
} catch (...) {
    hydro::set_exception_error(__func__);
    return HYDRO_ERR_NOMEM;
}
