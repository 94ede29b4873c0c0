/*
 * pasif.h -- C ABI of libpasif_hip.so, the MI355X (gfx950) implementation of
 * PASIf's batched RK4 forward integrator and its checkpointed adjoint.
 *
 * The entry points replace, one for one, the methods of the reference's
 * pybind11 class `__GpuDriver` (src/PASIfgpu.cpp:22-104, declared in
 * src/__GpuDriver.cuh:18-83).  Differences at the boundary:
 *   - every function returns an int status (PASIF_OK == 0); the message of
 *     the last failure is available from pasif_last_error().  The reference
 *     calls std::exit(0) on CUDA errors (src/helpers.cuh:43-51);
 *   - inputs are caller-owned buffers, copied before the call returns (the
 *     reference copies through pybind11's STL casters);
 *   - results are written into caller-owned buffers; pass out == NULL to
 *     query the length first.
 * Entry points marked "extension" have no counterpart in the reference.
 *
 * Indices of tensors are interleaved exactly as in the reference:
 * entry t of a 3-D tensor has indices idx[3t], idx[3t+1], idx[3t+2]
 * (output, source, set point; src/helpers.cu:277-282).
 */
#ifndef PASIF_H
#define PASIF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pasif_driver pasif_driver;

enum {
  PASIF_OK = 0,
  PASIF_ERR_ARGUMENT = 1,   /* invalid shape / value / order of calls      */
  PASIF_ERR_DEVICE = 2,     /* HIP runtime failure                          */
  PASIF_ERR_UNSUPPORTED = 3,/* valid input outside what the kernels handle  */
  PASIF_ERR_MEMORY = 4,     /* device memory / HBM budget exhausted         */
  PASIF_ERR_STATE = 5       /* call made before its prerequisites           */
};

enum { PASIF_SOLVER_PERSISTENT = 0, PASIF_SOLVER_PLAIN = 1, PASIF_SOLVER_GRAPH = 2 };
enum { PASIF_KERNEL_AUTO = 0, PASIF_KERNEL_TILE = 1, PASIF_KERNEL_LANE = 2 };
/* reported by pasif_kernel_info when several families share a solve */
enum { PASIF_KERNEL_MIXED = 3 };
/* row-parallel path: one workgroup per instance (extension) */
enum { PASIF_KERNEL_BIG = 4 };

/* __GpuDriver(excitationSet_, sampleRate_, numsteps_, dCompute_, dSystem_,
 * dSolver_, GPUId_)  (src/__GpuDriver.cu:25-157).  excitations is
 * [n_exc][exc_len] row-major.  numsteps == 0 means exc_len (PASIf.py:156-157
 * applies that default on the Python side). */
int pasif_create(pasif_driver** out, const double* excitations, uint32_t n_exc,
                 uint32_t exc_len, uint32_t sample_rate, uint32_t numsteps,
                 int d_compute, int d_system, int d_solver, int gpu_id);
/* ~__GpuDriver (src/__GpuDriver.cu:159-193) */
int pasif_destroy(pasif_driver* d);
/* message of the last failure on this handle ("" if none) */
const char* pasif_last_error(const pasif_driver* d);

/* _setFwdK / _setBwdK (src/__GpuDriver.cu:199-209, 272-282): CSR of the
 * scaled B+K, n = {rows, cols}, nnz values, nnz column indices, rows+1 ptr. */
int pasif_set_fwd_K(pasif_driver* d, const uint32_t n[2], const double* values,
                    const uint32_t* indices, const uint32_t* indptr, size_t nnz);
int pasif_set_bwd_K(pasif_driver* d, const uint32_t n[2], const double* values,
                    const uint32_t* indices, const uint32_t* indptr, size_t nnz);
/* _setFwdGamma / _setFwdLambda / _setFwdPsi and the Bwd versions
 * (src/__GpuDriver.cu:211-239, 284-312): order-3/4/5 COO tensors. */
int pasif_set_fwd_gamma(pasif_driver* d, const uint32_t n[3], const double* values,
                        const uint32_t* indices, size_t nnz);
int pasif_set_fwd_lambda(pasif_driver* d, const uint32_t n[4], const double* values,
                         const uint32_t* indices, size_t nnz);
int pasif_set_fwd_psi(pasif_driver* d, const uint32_t n[5], const double* values,
                      const uint32_t* indices, size_t nnz);
int pasif_set_bwd_gamma(pasif_driver* d, const uint32_t n[3], const double* values,
                        const uint32_t* indices, size_t nnz);
int pasif_set_bwd_lambda(pasif_driver* d, const uint32_t n[4], const double* values,
                         const uint32_t* indices, size_t nnz);
int pasif_set_bwd_psi(pasif_driver* d, const uint32_t n[5], const double* values,
                      const uint32_t* indices, size_t nnz);
/* _setFwdForcePattern / _setFwdInitialConditions and Bwd versions
 * (src/__GpuDriver.cu:242-254, 314-326).  Force entries with |v| <= eps of
 * the working precision are dropped (src/helpers.cu:674-680). */
int pasif_set_fwd_force_pattern(pasif_driver* d, const double* force, size_t n);
int pasif_set_fwd_initial_conditions(pasif_driver* d, const double* q0, size_t n);
int pasif_set_bwd_force_pattern(pasif_driver* d, const double* force, size_t n);
int pasif_set_bwd_initial_conditions(pasif_driver* d, const double* lam_T, size_t n);
/* _allocateSystemOnDevice / _allocateAdjointOnDevice
 * (src/__GpuDriver.cu:256-265, 328-337): validate, split the block-diagonal
 * global system into its independent systems, upload. */
int pasif_allocate_system(pasif_driver* d);
int pasif_allocate_adjoint(pasif_driver* d);

/* _loadExcitationsSet (src/__GpuDriver.cu:343-381) */
int pasif_load_excitations(pasif_driver* d, const double* excitations, uint32_t n_exc,
                           uint32_t exc_len, uint32_t sample_rate);
/* _setInterpolationMatrix (src/__GpuDriver.cu:383-398): [rows][window]
 * row-major, rows >= 2 (the kernels read rows 0 and 1).  size == 0 returns
 * to zero-order hold.  Any window >= 1 is accepted (the reference needs a
 * power of two). */
int pasif_set_interpolation_matrix(pasif_driver* d, const double* matrix, size_t size,
                                   uint32_t window);
/* _setModulationBuffer (src/__GpuDriver.cu:400-410): validated and stored;
 * no kernel reads it, as in the reference. */
int pasif_set_modulation_buffer(pasif_driver* d, const double* buffer, size_t n);

/* _getAmplitudes (src/__GpuDriver.cu:416-461): final state of every DOF of
 * every replica, [replica][dof].  *len receives the element count. */
int pasif_get_amplitudes(pasif_driver* d, double* out, size_t cap, size_t* len);
/* _getTrajectory (src/__GpuDriver.cu:463-500): floor(T/save_steps) slots of
 * the post-step state after each step t with t % save_steps == 0, laid out
 * [slot][replica][dof]. */
int pasif_get_trajectory(pasif_driver* d, uint32_t save_steps, double* out, size_t cap,
                         size_t* len);
/* _getGradient (src/__GpuDriver.cu:502-587): lambda_0 of the adjoint
 * system, [replica][adjoint dof], over the full schedule T..1.  `save` is
 * accepted for signature compatibility; nonzero values are rejected. */
int pasif_get_gradient(pasif_driver* d, uint32_t save, double* out, size_t cap, size_t* len);

/* ---------------- extensions (no counterpart in the reference) ---------- */
/* Per-system structure and scheduling.  sizes_fwd[s] / sizes_adj[s] give the
 * DOF count of system s in the block-diagonal global system (NULL:
 * detected from the sparsity).  exc_of_system[s] (NULL: every system runs
 * against every excitation, the reference's Cartesian product) maps each
 * system to one excitation.  steps_of_system[s] (NULL: numsteps) gives each
 * system its own step count. */
int pasif_set_systems_layout(pasif_driver* d, uint32_t n_systems, const uint32_t* sizes_fwd,
                             const uint32_t* sizes_adj, const int32_t* exc_of_system,
                             const uint32_t* steps_of_system);
/* 64 (default) or 32 */
int pasif_set_precision(pasif_driver* d, int bits);
/* PASIF_SOLVER_PERSISTENT (default): all steps in one launch.
 * PASIF_SOLVER_PLAIN: one launch per RK4 step, state through HBM.
 * PASIF_SOLVER_GRAPH: launches of at most `steps_per_launch` steps each,
 * captured `launches_per_graph` at a time into a hipGraph that is replayed
 * until every tile is done (bounded kernel durations). */
int pasif_set_solver(pasif_driver* d, int solver);
/* slicing of PASIF_SOLVER_GRAPH (defaults 256 steps, 16 launches) */
int pasif_set_graph_slicing(pasif_driver* d, uint32_t steps_per_launch,
                            uint32_t launches_per_graph);
/* tile lanes (64..1024, multiple of 64; 0 = automatic), slot-width
 * configuration (0 = automatic, 1 = narrow {4,4,2,2}, 2 = wide {8,8,4,2},
 * 3 = small {4,2,2,1} slots per lane for terms with 1..4 state factors) and
 * the adjoint checkpoint interval in steps (0 = automatic). */
int pasif_set_tuning(pasif_driver* d, uint32_t tile_lanes, uint32_t slot_config,
                     uint32_t checkpoint_every);
/* fused reduction: max |q_m| over the post-step states m in (T-window, T] of
 * every DOF; 0 disables.  Read with pasif_get_max_amplitudes after a solve. */
int pasif_set_amplitude_window(pasif_driver* d, uint32_t window);
int pasif_get_max_amplitudes(pasif_driver* d, double* out, size_t cap, size_t* len);
/* duration of the last solve's kernel launches (HIP events on the library's
 * stream), number of launches and number of tiles, lanes per tile */
int pasif_last_timing(const pasif_driver* d, double* kernel_ms, uint32_t* launches,
                      uint32_t* tiles, uint32_t* lanes);
/* Kernel family.  PASIF_KERNEL_TILE: rows of packed systems across the lanes
 * of a workgroup, stage values exchanged through LDS (systems of up to 1024
 * lanes of row segments).
 * PASIF_KERNEL_LANE: one lane per (system, excitation) with the state and
 * coefficients in registers, code generated per sparsity structure and
 * compiled at the first solve (small systems: PASIF_ERR_UNSUPPORTED when a
 * system does not fit; at most 8 distinct structures).  PASIF_KERNEL_AUTO
 * (default): the lane kernel for the structures that fit it (the 8 with the
 * most work), the tile kernel for the rest, launched side by side, and the
 * row-parallel path for systems that fit neither.  PASIF_KERNEL_BIG: every
 * instance on the row-parallel path -- one workgroup of 1024 threads per
 * (system, excitation), rows strided over its threads, state and stage
 * vectors in device memory, a barrier between stages; any system size, no
 * trajectory or amplitude window (PASIF_ERR_UNSUPPORTED), not for pool jobs
 * or pasif_solve_noalloc. */
int pasif_set_kernel(pasif_driver* d, int kernel);
/* what the packed problem runs on: kernel (TILE, LANE, BIG or MIXED), structure classes
 * (lane kernel; 0 for tiles) and the checkpoint interval the adjoint uses
 * (the lane kernel shortens it until a segment fits its LDS) */
int pasif_kernel_info(pasif_driver* d, int adjoint, int* kernel, uint32_t* classes,
                      uint32_t* checkpoint_every);
/* Lanes per instance in the lane kernel (extension).  0 (default): chosen per
 * structure class from a cost model (one lane per instance while the GPU's
 * SIMDs are busy; a group of 2..64 lanes, rows in bands across the group and
 * values exchanged by lane rotations, when the instances are too few to keep
 * them busy or a system does not fit one lane's registers).  A power of two
 * from 1 to 64 requests that group size for every class it fits. */
int pasif_set_lane_group(pasif_driver* d, uint32_t lanes);
/* Moves every class's group size `bias` steps along the sizes that fit it
 * (clamped), from the cost model's choice: +1 doubles the lanes per instance
 * where possible, -1 halves them.  Ignored when pasif_set_lane_group fixes a
 * size.  The autotuner probes -1, 0 and +1. (extension) */
int pasif_set_lane_bias(pasif_driver* d, int bias);
/* lanes per instance of each lane-kernel class of the packed problem
 * (out == NULL: count query) */
int pasif_lane_groups(pasif_driver* d, int adjoint, uint32_t* out, size_t cap, size_t* len);
/* Adjoint recursion (extension).  0 (default): the reference's recursion
 * (src/__GpuDriver.cu:746-783): RK4 on the adjoint system with set points
 * x_n, (x_n + x_{n-1})/2, x_{n-1} -- a continuous adjoint, within O(h^2) of
 * the derivative of the computed forward map.  1: the exact discrete adjoint
 * -- the reverse-mode derivative of the computed RK4 step, the adjoint system
 * evaluated at each step's forward stage inputs y4, y3, y2, y1 (recomputed
 * and stored per segment); it equals that derivative to round-off when the
 * adjoint system is the transposed Jacobian of the forward one.  Mode 1 runs
 * on the tile kernel (PASIF_ERR_UNSUPPORTED with PASIF_KERNEL_LANE). */
int pasif_set_adjoint_mode(pasif_driver* d, int mode);
/* Memory plan (extension): device bytes the run buffers and outputs of a
 * solve of the loaded problem take with the current settings (*run_bytes),
 * of which checkpoints (*ckpt_bytes), and the device memory a solve may use
 * now (*free_bytes: free memory plus the run buffers the handle already
 * holds; 0 on a host-only handle).  A solve whose run buffers exceed 80 % of
 * that is refused with PASIF_ERR_MEMORY before anything is allocated, as the
 * reference's planner keeps to 0.8 x free (src/__GpuDriver.cu:896-938). */
int pasif_run_footprint(pasif_driver* d, int adjoint, size_t* run_bytes, size_t* ckpt_bytes,
                        size_t* free_bytes);
/* Diagnostics: the generated HIP source of the lane kernel for the loaded
 * problem (NUL-terminated; out == NULL: length query), and with compile != 0
 * its hipRTC compilation for gfx950, which needs no GPU. */
int pasif_lane_source(pasif_driver* d, int adjoint, char* out, size_t cap, size_t* len,
                      int compile);
/* ---------------- continuous-batching pool (extension) ------------------
 * A pool keeps many requests ("jobs") resident on one GPU.  Each job is a
 * driver handle created by pasif_pool_job_create and loaded with the usual
 * setters (systems, layout, interpolation, adjoint); pasif_pool_submit packs
 * it into tiles of the pool's shape and admits it if the pool's HBM budget
 * allows (PASIF_ERR_MEMORY otherwise: retry after collecting finished jobs).
 * pasif_pool_tick advances every admitted, unfinished job by steps_per_tick
 * steps of work in ONE launch (wait = 0 returns after enqueueing, so the host
 * can admit new requests while the GPU runs; pasif_pool_wait joins).  A job's
 * results do not depend on which other jobs share the launches. */
typedef struct pasif_pool pasif_pool;
int pasif_pool_create(pasif_pool** out, int gpu_id, int precision, uint32_t tile_lanes,
                      uint32_t slot_config, uint32_t checkpoint_every, uint32_t steps_per_tick,
                      double hbm_budget_bytes);
/* Destroys the pool.  Job handles it created stay with their owners: jobs
 * still resident lose their device memory and every such handle is detached
 * from the pool (pasif_destroy it as usual).  pasif_destroy on a job handle
 * removes the job from its pool first (a job a tick in flight still runs is
 * deleted by the pool once that tick is over). */
int pasif_pool_destroy(pasif_pool* pool);
const char* pasif_pool_last_error(const pasif_pool* pool);
/* Signal bank shared by jobs created with excitations == NULL: [n_exc][len]
 * samples, the sample rate, and optionally interpolation rows 0 and 1
 * ([2][window]).  Uploaded once. */
int pasif_pool_set_bank(pasif_pool* pool, const double* excitations, uint32_t n_exc,
                        uint32_t exc_len, uint32_t sample_rate, const double* interp,
                        uint32_t window);
/* a job handle (a driver on the pool's device); excitations == NULL uses the
 * pool's signal bank (excitation ids index the bank) */
int pasif_pool_job_create(pasif_pool* pool, pasif_driver** job, const double* excitations,
                          uint32_t n_exc, uint32_t exc_len, uint32_t sample_rate,
                          uint32_t numsteps);
/* Packs and admits a loaded job.  A system of a job on the pool's signal
 * bank whose sparsity structure a lane module of the pool runs (a
 * registered structure, any coefficient values) runs on that module; the
 * job's other systems, and every system of a job with its own excitations,
 * run on tiles. */
int pasif_pool_submit(pasif_pool* pool, pasif_driver* job, int adjoint);
int pasif_pool_tick(pasif_pool* pool, int wait, uint32_t* active_tiles);
int pasif_pool_wait(pasif_pool* pool, double* kernel_ms);
int pasif_pool_job_finished(pasif_pool* pool, const pasif_driver* job, int* finished);
/* final states [dof] and, for adjoint jobs, lambda_0 [adjoint dof]; frees the
 * job's device memory and removes it from the pool (then pasif_destroy it).
 * A job finished by an earlier tick is left out of later ticks and may be
 * collected while one runs; otherwise collect after pasif_pool_wait. */
int pasif_pool_collect(pasif_pool* pool, pasif_driver* job, double* q, size_t capq, size_t* lenq,
                       double* grad, size_t capg, size_t* leng);
int pasif_pool_stats(const pasif_pool* pool, uint32_t* jobs, double* used_bytes,
                     double* last_kernel_ms, uint32_t* last_blocks);
/* Slot cap: at most max_tiles tiles and lane workgroups that ticks still
 * launch (0 = no cap; the HBM budget still applies).  A lane workgroup
 * leaves the count once the completed ticks cover the work of its longest
 * system, a finished job's tiles once it has finished (its memory stays
 * with it until it is collected).  A submit that would exceed the cap
 * returns PASIF_ERR_MEMORY like an exhausted budget (retry after a tick or
 * a collect).  A job is always admitted into an empty pool. */
int pasif_pool_set_slots(pasif_pool* pool, uint32_t max_tiles);
/* slots held now (tiles and lane workgroups still launched) and the cap */
int pasif_pool_slots(const pasif_pool* pool, uint32_t* used, uint32_t* max_tiles);
/* on != 0: structures of jobs carrying their own arrays (on the signal bank)
 * that no lane module runs are compiled into further lane modules on a
 * thread of the pool's own (half the queue per module, 4 to 16 classes); a
 * finished module is taken in by the next submit or tick, and later jobs of
 * those structures run on it (jobs admitted before stay on tiles).  Each
 * module is one more launch per tick, on a stream of its own: once nothing
 * is queued and the single-wave classes compiled so far span three or more
 * modules, they are compiled again as one module, which later jobs use.
 * Off by default. */
int pasif_pool_set_compile(pasif_pool* pool, int on);
/* Waits for at most max_builds module compilations (0: until every queued
 * structure has its module or failed to compile; those stay on tiles), each
 * taken in when done: modules built and failed so far, structures still
 * queued or compiling. */
int pasif_pool_compile_wait(pasif_pool* pool, uint32_t max_builds, uint32_t* modules,
                            uint32_t* failed, uint32_t* queued);
/* Register every system of `src` (a handle, host-only or not, whose system
 * -- and, optionally, adjoint -- was allocated) with the pool: ids
 * first_id .. first_id + count - 1 in the handle's system order.  Requests
 * then name systems by id and only their excitation ids and step counts
 * travel per request. */
int pasif_pool_register(pasif_pool* pool, const pasif_driver* src, uint32_t* first_id,
                        uint32_t* count);
/* Create, pack and admit a job of n registered systems in one call: system
 * i is ids[i] driven by bank excitation exc_of_system[i] (-1: unforced) for
 * steps_of_system[i] steps.  On PASIF_ERR_MEMORY (no slots / budget) *job is
 * NULL and nothing is held. */
int pasif_pool_submit_systems(pasif_pool* pool, uint32_t n, const uint32_t* ids,
                              const int32_t* exc_of_system, const uint32_t* steps_of_system,
                              int adjoint, pasif_driver** job);
/* Many requests in one call: request r is systems sys_off[r] ..
 * sys_off[r+1]-1 of ids / exc_of_system / steps_of_system, forward + adjoint
 * when adjoint[r] != 0.  Requests are admitted in order until one does not
 * fit the slots or the HBM budget; *admitted counts those admitted (their
 * handles in jobs[0 .. *admitted-1], the rest NULL) and the call still
 * returns PASIF_OK.  Any other failure returns its code, with the requests
 * before it admitted. */
int pasif_pool_submit_batch(pasif_pool* pool, uint32_t n_req, const uint32_t* sys_off,
                            const uint32_t* ids, const int32_t* exc_of_system,
                            const uint32_t* steps_of_system, const int32_t* adjoint,
                            pasif_driver** jobs, uint32_t* admitted);
/* pasif_pool_collect for n finished jobs, then pasif_destroy of each handle:
 * job i's final states go to q[q_off[i] .. q_off[i+1]) (exactly that many),
 * its lambda_0 (adjoint jobs) to grad[g_off[i] .. g_off[i+1]); grad may be
 * NULL when no job is an adjoint job. */
int pasif_pool_collect_batch(pasif_pool* pool, uint32_t n, pasif_driver* const* jobs, double* q,
                             const size_t* q_off, double* grad, const size_t* g_off);
/* *running = 1 while the last tick's launch is still executing */
int pasif_pool_poll(pasif_pool* pool, int* running);

/* ---------------- no-allocation solve (extension) -----------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the forward solve (adjoint == 0, as pasif_get_amplitudes)
 * or the forward + adjoint solve (adjoint != 0, as pasif_get_gradient) of the
 * problem already loaded in `d`, with every run buffer (states, progress
 * records, checkpoints, recomputed segments, kernel argument tables) in a
 * static device array of PASIF_NOALLOC_SCRATCH_BYTES per GPU, held by the
 * library until the process ends.  It calls no hipMalloc / hipFree.
 *  - Set-up (allocating, once per problem and settings): the output-length
 *    query pasif_get_amplitudes(d, NULL, 0, &n) or pasif_get_gradient(d, 0,
 *    NULL, 0, &n) packs the problem and uploads it and the excitations.
 *  - out_q: device buffer of the final states (pasif_get_amplitudes' length),
 *    out_grad: device buffer of lambda_0 (pasif_get_gradient's length,
 *    adjoint only), both of the working precision (double for 64 bits,
 *    float for 32).  Results equal those of the allocating entry points.
 *  - Returns when the results are in place.  PASIF_ERR_MEMORY if the run
 *    buffers exceed the scratch array, PASIF_ERR_STATE without the set-up,
 *    PASIF_ERR_UNSUPPORTED with an amplitude window, the graph solver or a
 *    pool job.
 *  - Every call on one GPU shares that GPU's array: calls that overlap in
 *    time on the same GPU (two threads, two drivers) race and must be
 *    serialised by the caller. */
#define PASIF_NOALLOC_SCRATCH_BYTES (64u << 20)
int pasif_solve_noalloc(pasif_driver* d, int adjoint, void* out_q, void* out_grad);

/* library build string */
const char* pasif_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PASIF_H */
