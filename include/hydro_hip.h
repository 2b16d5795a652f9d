/*
 * hydro_hip.h -- C ABI of the MI355X (gfx950) first-order finite-volume Euler solver.
 *
 * Plain C: pointers, sizes and an opaque stream (a hipStream_t passed as void*, NULL =
 * the default stream).  Every device pointer is HIP device memory laid out as in the
 * reference (hydro2d/main.cxx:118-128, hydro3d/main.cxx:122-133):
 *   2D: cell n = i*(NB_Y+2) + j;          rhoE[2n] = rho, rhoE[2n+1] = E,
 *       uv[2n] = rho*u, uv[2n+1] = rho*v
 *   3D: cell n = i*(NB_Y+2)*(NB_Z+2) + j*(NB_Z+2) + k;  rhoE as in 2D,
 *       uvw[3n+{0,1,2}] = rho*u, rho*v, rho*w
 * with one ghost layer on every side; NB_* are interior sizes.
 *
 * Errors: functions returning int give 0 on success and a negative code on failure;
 * the launcher-style functions returning double give the elapsed microseconds, or a
 * negative value on failure.  hydro_last_error() describes the last failure of the
 * calling thread.
 */
#ifndef HYDRO_HIP_H_
#define HYDRO_HIP_H_

#include <stddef.h>

#define HYDRO_HIP_ABI_VERSION 1

/* Boundary modes (the reference hard-wires X periodic, Y and Z reflective:
 * hydro2d/main.cxx:349-378 and hydro2d/kernel.cxx:213-251). */
#define HYDRO_BC_PERIODIC 0
#define HYDRO_BC_REFLECTIVE 1

/* State / arithmetic precision of the run and batch APIs. */
#define HYDRO_FP64 64      /* the reference scheme in FP64 (bit-identical) */
#define HYDRO_FP32 32      /* the reference's USE_FLOAT build (bit-identical) */
#define HYDRO_FP32_FLUX 3264 /* FP64 state; FP32 primitives, fluxes and signal speeds with
                              * hardware reciprocal / square root; increments, update, Dt and
                              * reductions in FP64 (compared to FP64 with a tolerance) */

#define HYDRO_ERR_ARG (-1)
#define HYDRO_ERR_HIP (-2)
#define HYDRO_ERR_NONFINITE (-3)
#define HYDRO_ERR_NOMEM (-4)

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------
 * Drop-in launchers.
 *
 * Concurrency (launcher*, hydro_step2d, hydro_step3d): steps on different streams, or
 * from different host threads, may overlap on the device.  Each stream of each device
 * (each thread, for hipStreamPerThread) has its own 512-byte Dt reduction state, bound
 * to it at its first step and kept until the process ends.  The states come from a pool
 * of 256 per device in the library's device image, so the launchers allocate nothing;
 * while all 256 are bound, a further stream gets a hipMalloc'd state.  Steps on one
 * stream run in stream order, as any HIP work does.  hydro_stream_release unbinds a
 * stream's state (before the stream is destroyed, when its handle may be reused or
 * many streams come and go); without it a state stays bound until the process ends,
 * and a new stream that reuses the handle of a destroyed one takes over its state.
 * ------------------------------------------------------------------------------- */

/* Release the Dt reduction state bound to `stream` on the current device (for
 * hipStreamPerThread: the calling thread's).  Waits for the stream, then returns the
 * state to the device's pool, or frees it if it was allocated.  No effect if no step ran
 * on the stream.  No counterpart in the reference (its launcher owns one queue). */
int hydro_stream_release(void* stream);

/* Number of streams with a bound state on the current device, and of the device's 256
 * pooled states that are free (diagnostics for hydro_stream_release). */
int hydro_stream_states(int* bound, int* pooled_free);

/* Replaces `launcher` of hydro2d/kernel.hpp:63-68 (defined at hydro2d/kernel.cxx:331-367).
 * Same parameter order and meaning; sycl::queue becomes a hipStream_t and the
 * `const DATATYPE&` ratios are passed by value.  Reads d_rhoE/d_uv including the X
 * ghost planes the caller filled, writes only the interior cells of d_*_next and
 * *Dt_next = C_min / max signal speed of the new state, with reflective walls at
 * j = 1 and j = NB_Y.  Synchronous like the reference (it waits for the kernel) and
 * returns the kernel's elapsed microseconds (HIP events). */
double launcher(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next, double* Dt_next,
                double C_min, double gamma, double gamma_minus_one, double divgamma, double K, size_t NB_X,
                size_t NB_Y, double DtDx, double DtDy, void* stream);

/* Replaces `launcher` of the 3D build, hydro3d/kernel.hpp:64-69 (hydro3d/kernel.cxx:
 * 419-455), renamed so that one library carries both dimensions.  Reflective walls at
 * j = 1, NB_Y and k = 1, NB_Z. */
double launcher3d(const double* d_rhoE, const double* d_uvw, double* d_rhoE_next, double* d_uvw_next,
                  double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma, double K,
                  size_t NB_X, size_t NB_Y, size_t NB_Z, double DtDx, double DtDy, double DtDz, void* stream);

/* The same two launchers for the USE_FLOAT build (hydro2d/kernel.hpp:9-23,
 * hydro2d/Makefile:7-9): every state value and constant is float. */
double launcher_f32(const float* d_rhoE, const float* d_uv, float* d_rhoE_next, float* d_uv_next, float* Dt_next,
                    float C_min, float gamma, float gamma_minus_one, float divgamma, float K, size_t NB_X,
                    size_t NB_Y, float DtDx, float DtDy, void* stream);
double launcher3d_f32(const float* d_rhoE, const float* d_uvw, float* d_rhoE_next, float* d_uvw_next, float* Dt_next,
                      float C_min, float gamma, float gamma_minus_one, float divgamma, float K, size_t NB_X,
                      size_t NB_Y, size_t NB_Z, float DtDx, float DtDy, float DtDz, void* stream);

/* ---------------------------------------------------------------------------------
 * Asynchronous forms with boundary modes (no wait, no timing).
 * ------------------------------------------------------------------------------- */

/* One step as `launcher`, enqueued on `stream`.  bc_y selects reflective walls
 * (HYDRO_BC_REFLECTIVE, the reference) or reading the Y ghost cells
 * (HYDRO_BC_PERIODIC, filled by hydro_ghost_fill2d). */
int hydro_step2d(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next, double* Dt_next,
                 double C_min, double gamma, double gamma_minus_one, double divgamma, double K, size_t NB_X,
                 size_t NB_Y, double DtDx, double DtDy, int bc_y, void* stream);
int hydro_step3d(const double* d_rhoE, const double* d_uvw, double* d_rhoE_next, double* d_uvw_next,
                 double* Dt_next, double C_min, double gamma, double gamma_minus_one, double divgamma, double K,
                 size_t NB_X, size_t NB_Y, size_t NB_Z, double DtDx, double DtDy, double DtDz, int bc_y, int bc_z,
                 void* stream);

/* Ghost fill on the device.  Replaces update_ghost_cells_single_device
 * (hydro2d/main.cxx:349-378, hydro3d/main.cxx:374-405) together with the host->device
 * ghost-plane copies (hydro2d/main.cxx:719-729) and, for a domain resident on one GPU,
 * the MPI halo exchange (hydro2d/main.cxx:381-424).
 *   bc_x PERIODIC:   plane 0 <- plane NB_X, plane NB_X+1 <- plane 1 (all j, k).
 *   bc_x REFLECTIVE: plane 0 <- mirror(plane 1), plane NB_X+1 <- mirror(plane NB_X),
 *                    rho*u negated.
 *   bc_y / bc_z PERIODIC: Y (Z) ghost lines <- the opposite interior line, in every
 *                    plane, after the X planes.  REFLECTIVE: Y (Z) ghosts untouched
 *                    (the walls are applied inside the step kernel). */
int hydro_ghost_fill2d(double* d_rhoE, double* d_uv, size_t NB_X, size_t NB_Y, int bc_x, int bc_y, void* stream);
int hydro_ghost_fill3d(double* d_rhoE, double* d_uvw, size_t NB_X, size_t NB_Y, size_t NB_Z, int bc_x, int bc_y,
                       int bc_z, void* stream);
int hydro_ghost_fill2d_f32(float* d_rhoE, float* d_uv, size_t NB_X, size_t NB_Y, int bc_x, int bc_y, void* stream);
int hydro_ghost_fill3d_f32(float* d_rhoE, float* d_uvw, size_t NB_X, size_t NB_Y, size_t NB_Z, int bc_x, int bc_y,
                           int bc_z, void* stream);

/* ---------------------------------------------------------------------------------
 * Device-resident step loop: the host loop of hydro2d/main.cxx:686-815 with Dt,
 * t and the iteration count kept on the device, so that steps enqueue back to back
 * with no host round trip.  Results are identical to calling hydro_ghost_fill* and
 * launcher* per step and advancing t on the host.
 * ------------------------------------------------------------------------------- */
typedef struct hydro_run hydro_run;

/* dim = 2 or 3; nb[dim] interior sizes; D[dim] the cell spacings DX, DY(, DZ) as the
 * reference computes them (ONE / (NB + 1), hydro2d/main.cxx:482-483); C_min = C *
 * min(D); dt0 the initial Dt; the loop runs while t < max_t and it < max_it.  The
 * per-iteration (t, Dt) pairs of up to max_it steps are recorded on the device.
 * precision: HYDRO_FP64, HYDRO_FP32 or HYDRO_FP32_FLUX.  Returns NULL on error. */
hydro_run* hydro_run_create(int dim, const size_t* nb, const double* D, double C_min, double dt0, double max_t,
                            size_t max_it, int precision, const int* bc);

/* Enqueue up to n_steps steps on `stream`, alternating between the (a) and (b) state
 * buffers starting from whichever holds the current state (a before the first step).
 * Each step fills ghosts then updates.  Steps after the loop ends are no-ops. */
int hydro_run_advance(hydro_run* run, void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b, size_t n_steps,
                      void* stream);

/* Wait for `stream` and report the loop state: completed iterations, t, the current Dt,
 * whether the loop has ended, and which buffer holds the current state (0 = a, 1 = b).
 * Returns HYDRO_ERR_NONFINITE if a Dt came out NaN or infinite. */
int hydro_run_status(hydro_run* run, size_t* it, double* t, double* dt, int* done, int* current, void* stream);

/* Copy recorded (t, Dt) pairs [from, from+count) to host arrays (either may be NULL). */
int hydro_run_trajectory(hydro_run* run, size_t from, size_t count, double* t_out, double* dt_out, void* stream);

/* Record HIP events around every step-kernel launch of later hydro_run_advance calls
 * (enable != 0), for measuring the kernel alone.  At most 65536 launches are recorded
 * between two hydro_run_kernel_time calls; later ones run untimed. */
int hydro_run_set_timing(hydro_run* run, int enable);

/* Wait for the recorded launches and return the sum of their kernel times in ms and
 * their count; clears the record. */
int hydro_run_kernel_time(hydro_run* run, double* total_ms, size_t* launches);

void hydro_run_destroy(hydro_run* run);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc; it waits for its stream, so it
 * is not for use under graph capture): the resident loop of hydro_run_create +
 * hydro_run_advance + hydro_run_status + hydro_run_trajectory in one synchronous call
 * that calls no hipMalloc, hipFree or other allocating routine.  Arguments are those of
 * hydro_run_create and hydro_run_advance; the call runs the loop to its end (at most
 * max_it steps from buffer a), waits for `stream`, and reports what hydro_run_status
 * reports (it, t, the current Dt, the buffer holding the current state: 0 = a, 1 = b)
 * and, when traj_t / traj_dt are not NULL, the it recorded (t, Dt) pairs into those host
 * arrays.  The results are bit-identical to the allocating entry points.
 * Its loop state lives in a static device array of the library (1.25 MiB, one per device
 * used, kept until the process ends), which limits max_it to HYDRO_NOALLOC_MAX_IT: a
 * larger max_it is rejected with HYDRO_ERR_ARG.  Every call on a device shares that
 * device's array: calls that overlap on one device (two host threads, or two streams)
 * race and must not be made; successive calls are independent (each call resets the
 * state it uses).  Returns 0, HYDRO_ERR_NONFINITE as hydro_run_status, or an error. */
#define HYDRO_NOALLOC_MAX_IT 65536
int hydro_run_noalloc(int dim, const size_t* nb, const double* D, double C_min, double dt0, double max_t,
                      size_t max_it, int precision, const int* bc, void* rhoE_a, void* mom_a, void* rhoE_b,
                      void* mom_b, size_t* it, double* t, double* dt, int* current, double* traj_t, double* traj_dt,
                      void* stream);

/* Tuning knobs of the launch geometry (no effect on results), process-wide:
 *   "rows2d" = maximum X rows per wavefront strip of the 2D kernel (4..4096);
 *   "occ2d"  = workgroups per CU of the 2D kernel (3..8, enforced by reserving LDS;
 *              8 leaves the count to the kernel's registers);
 *   "rows3d" = maximum X rows per workgroup strip of the 3D kernel (2..4096).
 * 0 selects the per-precision defaults: 2D 12 rows with 3 workgroups per CU (FP64 and
 * FP32-flux), 16 rows with 6 for FP32; 3D 32 rows.  Grids with few workgroups
 * use shorter strips than the maximum.  hydro_tune_get returns a knob's setting, or with
 * the suffix "_last" ("rows2d_last", "occ2d_last", "rows3d_last") the geometry of the
 * last step launched.  Returns 0, or HYDRO_ERR_ARG for an unknown key or out-of-range
 * value. */
int hydro_tune_set(const char* key, int value);
int hydro_tune_get(const char* key);

/* ---------------------------------------------------------------------------------
 * Initial condition.
 * ------------------------------------------------------------------------------- */

/* Blast initial condition on the device.  Replaces set_init_conditions_diffusion
 * (hydro2d/main.cxx:280-338, hydro3d/main.cxx:296-363) plus the host->device copy of the
 * initial state (hydro2d/main.cxx:650-652).  Writes the interior cells of a (sub)domain
 * of interior size nb[dim] whose first interior cell has global 0-based index base[dim]
 * (NULL = origin) in a global grid NB[dim]; ghost cells are not touched.  rho = rho_in /
 * rho_out, E = p_in / (gamma-1) or p_out / (gamma-1), zero momentum.  ic_mode 0: the
 * reference disc / sphere (centre (NB+1)/2, radius min(centre)/3); ic_mode 1: radius =
 * radius_frac * min(NB + 1).  *min_dt receives min over cells of C_min / (c_s + |u|). */
int hydro_ic(double* d_rhoE, double* d_mom, int dim, const size_t* nb, const size_t* base, const size_t* NB,
             double rho_in, double rho_out, double p_in, double p_out, int ic_mode, double radius_frac, double C_min,
             double gamma, double* min_dt, void* stream);
int hydro_ic_f32(float* d_rhoE, float* d_mom, int dim, const size_t* nb, const size_t* base, const size_t* NB,
                 float rho_in, float rho_out, float p_in, float p_out, int ic_mode, double radius_frac, float C_min,
                 float gamma, float* min_dt, void* stream);

/* ---------------------------------------------------------------------------------
 * Batched form: many independent solves of uneven sizes in one engine.
 *
 * No single counterpart in the reference, which runs one solve per process; each blast
 * of a batch is one complete run of hydro2d/main.cxx:686-815 (hydro3d/main.cxx:742-876)
 * with the reference's boundaries (X periodic, Y and Z reflective) and step arithmetic
 * (hydro2d/kernel.cxx:132-319, hydro3d/kernel.cxx:145-407), bit-identical to a
 * single-domain hydro_run of the same blast.  One kernel launch per step advances
 * every active blast (a persistent grid walking a segment table, or one wave per tile
 * dispatched in table order; see hydro_batch_set_tiling), followed by one finalise launch that computes each blast's Dt,
 * advances its t / iteration count and flips its current buffer; a blast stops on its
 * own when !(t < max_t && it < max_it).  Steps may be replayed from a hipGraph.  The
 * batch runs on its own non-blocking stream (hydro_batch_stream): work queued on a
 * blast's buffers in other streams must be complete, or ordered before that stream
 * (hipStreamWaitEvent), when hydro_batch_add / hydro_batch_add_blast is called, since
 * add_blast writes the initial state on the batch stream at once.  A batch is used
 * from one host thread at a time; distinct batches may run concurrently.
 * ------------------------------------------------------------------------------- */

typedef struct hydro_batch hydro_batch;

/* dim 2 or 3; precision HYDRO_FP64 or HYDRO_FP32_FLUX (FP64 state); capacity = maximum
 * number of blasts held at once.  Uses the current device.  NULL on failure. */
hydro_batch* hydro_batch_create(int dim, int precision, size_t capacity);
void hydro_batch_destroy(hydro_batch* batch);
void* hydro_batch_stream(hydro_batch* batch);

/* Add a blast whose state (reference layout, interior sizes nb[dim]) is in rhoE_a/mom_a;
 * rhoE_b/mom_b are its second buffer pair.  D = grid spacings, dt0 = initial Dt (as
 * hydro_run_create).  Returns the slot (>= 0) or a negative error code. */
int hydro_batch_add(hydro_batch* batch, const size_t* nb, const double* D, double C_min, double dt0, double max_t,
                    size_t max_it, void* rhoE_a, void* mom_a, void* rhoE_b, void* mom_b);

/* Add a blast and write its initial condition on the device (asynchronously, on the
 * batch stream): hydro_ic's parameterised blast (ic_mode 1: radius = radius_frac *
 * min(nb + 1), centre (nb + 1) / 2), spacings D = 1 / (nb + 1), C_min = C * min(D)
 * (hydro2d/main.cxx:482-484, 595-607), initial Dt reduced on the device. */
int hydro_batch_add_blast(hydro_batch* batch, const size_t* nb, double C, double radius_frac, double rho_in,
                          double rho_out, double p_in, double p_out, double max_t, size_t max_it, void* rhoE_a,
                          void* mom_a, void* rhoE_b, void* mom_b);

/* Free a slot (waits for the batch stream; the buffers stay the caller's). */
int hydro_batch_remove(hydro_batch* batch, int slot);

/* Enqueue n_steps steps of every active blast (asynchronous). */
int hydro_batch_advance(hydro_batch* batch, size_t n_steps);
int hydro_batch_sync(hydro_batch* batch);

/* Iteration, t, next Dt, stop flag and current buffer (0 = a, 1 = b) of one slot (waits
 * for the batch stream).  Returns HYDRO_ERR_NONFINITE if its Dt became NaN or inf. */
int hydro_batch_status(hydro_batch* batch, int slot, size_t* it, double* t, double* dt, int* done, int* current);

/* Stop flags of slots [0, n) in one copy: 0 running, 1 done, 2 stopped on a
 * non-finite Dt (waits for the batch stream). */
int hydro_batch_done_flags(hydro_batch* batch, unsigned char* flags, size_t n);

/* Tiling: workgroups per CU (0 = default), tile rows (0 = balanced: equal contiguous
 * row ranges per worker of a persistent grid; R > 0: tiles of R rows in row-major order
 * within each blast) and, in 2D, tile width in 62-column windows (1, 2 or 4 adjacent
 * windows marched together by one worker of as many waves; 0 = 1).
 *   Persistent grid (balanced, multi-window tiles, 3D): blocks_per_cu sizes the grid
 *   and the tiles are dealt round-robin to its workers.
 *   Direct dispatch (2D, R > 0, one window): one wave per tile in table order;
 *   blocks_per_cu caps the resident workgroups per CU (values below 3 act as 3; 0 = as
 *   registers allow), and R is halved, down to 8, until the step has at least 2048
 *   workgroups.
 * Results do not depend on the tiling. */
int hydro_batch_set_tiling(hydro_batch* batch, int blocks_per_cu, int tile_rows, int tile_windows);
int hydro_batch_get_tiling(hydro_batch* batch, int* blocks_per_cu, int* tile_rows, int* tile_windows);

/* Replay steps from a hipGraph of steps_per_graph steps (0 = direct launches). */
int hydro_batch_set_graph(hydro_batch* batch, size_t steps_per_graph);

/* Time every tiling candidate for steps_per_trial steps on the batch itself (they are
 * ordinary steps of the solves) and keep the fastest; with use_cache, a batch of the
 * same signature (dimension, precision, blast count, total cells) tuned earlier in the
 * process reuses that result without trials (*best_ms_per_step = -1 then). */
int hydro_batch_autotune(hydro_batch* batch, size_t steps_per_trial, int use_cache, double* best_ms_per_step);

/* ---------------------------------------------------------------------------------
 * Served form: a request loop with continuous batching over the batched engines.
 *
 * No counterpart in the reference (one solve per process run).  Each request is one
 * blast solve (as hydro_batch_add_blast).  Requests queue on submission and are admitted
 * in arrival order as soon as their device footprint -- two FP64 buffer pairs with
 * ghosts, 256-byte aligned -- fits in the server's HBM arena (one allocation made at
 * create) and their engine has a free slot.  One engine per (dimension, precision), each
 * on its own stream; a request joins between ticks and leaves when its own loop ends.
 * hydro_server_tick serves every idle engine (retire, admit, enqueue steps_per_tick
 * steps) and waits only while every engine with work is busy.  Times are seconds on the
 * server's steady clock (hydro_server_now).  A server is used from one host thread.
 * ------------------------------------------------------------------------------- */

typedef struct hydro_server hydro_server;

/* hbm_bytes: size of the state arena; capacity: running requests per engine. */
hydro_server* hydro_server_create(size_t hbm_bytes, size_t capacity, size_t steps_per_tick);
void hydro_server_destroy(hydro_server* server);
double hydro_server_now(hydro_server* server);

/* Queue a request; precision HYDRO_FP64 or HYDRO_FP32_FLUX; keep_result != 0 keeps the
 * final state in the arena until hydro_server_release (else it is freed on completion);
 * arrival = its arrival time on the server clock (< 0: now).  Returns the request id (>= 0)
 * or a negative error (HYDRO_ERR_NOMEM if it can never fit). */
int hydro_server_submit(hydro_server* server, int dim, const size_t* nb, int precision, double C, double radius_frac,
                        double rho_in, double rho_out, double p_in, double p_out, double max_t, size_t max_it,
                        int keep_result, double arrival);

/* One tick; *n_done = requests completed in it. */
int hydro_server_tick(hydro_server* server, size_t* n_done);
int hydro_server_pending(hydro_server* server, size_t* queued, size_t* running, size_t* bytes_in_use);

/* state: 0 queued, 1 running, 2 done (result kept), 3 done (memory released), 4 failed
 * (non-finite Dt).  Returns the request's error code (0 unless failed). */
int hydro_server_request(hydro_server* server, int id, int* state, double* t_arrival, double* t_admit,
                         double* t_done, size_t* it, double* t_sim);

/* Copy a kept result (current buffer pair, reference layout, ghost cells undefined) to
 * host or device memory; then release it. */
int hydro_server_result(hydro_server* server, int id, void* dst_rhoE, void* dst_mom);
int hydro_server_release(hydro_server* server, int id);
int hydro_server_stats(hydro_server* server, size_t* ticks, double* cell_updates);

/* Admission cap on the interior cells of the running requests of each engine (0 =
 * memory and slots only; a request is always admitted into an idle engine) and the
 * steps per tick.  A cap keeps large requests of one kind closer to first-come
 * first-served instead of sharing their engine with every later arrival. */
int hydro_server_set_limits(hydro_server* server, double max_active_cells, size_t steps_per_tick);

/* Size classes: a request whose estimated solve time alone (max_it steps of the plain
 * resident loop, each a launch floor plus a cost per interior cell measured on MI355X)
 * is at least large_seconds is "large".  Large requests run one at a time, in arrival
 * order, as plain device-resident solves on a stream of their own, their steps enqueued
 * by hydro_server_tick about 2 ms of estimated work at a time; smaller requests are
 * admitted to the batched engines beside them at once.  0 (the default) puts every
 * request in one class.  Applies to requests submitted after the call. */
int hydro_server_set_policy(hydro_server* server, double large_seconds);

/* Backlog of the large class: while more than large_backlog large requests wait, the
 * oldest waiting ones are admitted to their batched engines as small ones are (under a
 * long queue, one plain solve at a time is the slowest way through it).  0 (the
 * default) keeps every large request in the one-at-a-time order. */
int hydro_server_set_backlog(hydro_server* server, size_t large_backlog);

/* ---------------------------------------------------------------------------------
 * Diagnostics.
 * ------------------------------------------------------------------------------- */

/* Total mass (sum of interior rho) computed on the device with a compensated FP64
 * reduction; the reference computes it on the host with a compensated long double sum
 * (hydro2d/main.cxx:237-265).  nb_z = 0 for 2D.  precision 64 or 32. */
int hydro_total_mass(const void* d_rhoE, size_t NB_X, size_t NB_Y, size_t NB_Z, int precision, double* mass,
                     void* stream);

/* Total energy (sum of interior E), the same compensated FP64 reduction.  No counterpart
 * in the reference, which tracks mass only; it is the second conserved total the
 * FP32-flux form is checked on (periodic X and reflective walls conserve it exactly in
 * exact arithmetic). */
int hydro_total_energy(const void* d_rhoE, size_t NB_X, size_t NB_Y, size_t NB_Z, int precision, double* energy,
                       void* stream);

/* Copy `bytes` (a multiple of 16) from src to dst on the device with a plain streaming
 * kernel: the achievable-bandwidth reference for the step kernels, which read and write
 * the state once.  No counterpart in the reference (measurement only). */
int hydro_stream_copy(void* dst, const void* src, size_t bytes, void* stream);

/* The same copy with non-temporal loads and stores: the fastest read-once / write-once
 * stream of this library, quoted as the achievable ceiling of the step kernels.  No
 * counterpart in the reference (measurement only). */
int hydro_stream_copy_nt(void* dst, const void* src, size_t bytes, void* stream);

/* The 2D FP64 step's memory schedule without its arithmetic: the same workgroups,
 * 62-column windows, strips (the "rows2d" and "occ2d" knobs and their defaults), block
 * order and rows of loads in flight as the step kernel, loading every row the step loads
 * (halo rows included, the X ghost planes read directly) and storing every owned cell:
 * each interior cell of (d_rhoE, d_uv) is copied unchanged to (d_rhoE_next, d_uv_next);
 * ghost cells of the destination are not written.  Its rate is the ceiling of the step's
 * access pattern.  Layout and sizes as `launcher`.  No counterpart in the reference
 * (measurement only). */
int hydro_stream_pattern2d(const double* d_rhoE, const double* d_uv, double* d_rhoE_next, double* d_uv_next,
                           size_t NB_X, size_t NB_Y, void* stream);

/* Read-only and write-only streams over `bytes` (a multiple of 16): the achievable
 * bandwidth of each direction alone.  hydro_stream_read XOR-folds every 64-bit word of
 * src into the 256 slots of d_xor256 (device memory the caller zeroes first; the XOR of
 * the 256 slots is the XOR of all words).  hydro_stream_fill writes `pattern` to every
 * 64-bit word of dst.  No counterpart in the reference (measurement only). */
int hydro_stream_read(const void* src, size_t bytes, unsigned long long* d_xor256, void* stream);
int hydro_stream_fill(void* dst, size_t bytes, unsigned long long pattern, void* stream);

const char* hydro_last_error(void);
int hydro_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HYDRO_HIP_H_ */
