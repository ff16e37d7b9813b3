/*
 * pfs.h — C ABI of the MI355X (gfx950) fluid-step library libpfs.so.
 *
 * The library replaces the reference's solver implementations
 * (src/fluid.cpp and src/fluid.cu) behind the functions declared in the
 * reference's includes/fluid.hpp.  Results follow src/fluid.cpp, the CPU
 * path, bit for bit: same operation order, same buffer ping-pong, same stale
 * channels, same even/odd Jacobi-count behaviour.
 *
 * Conventions
 *   - Plain C types only.  A field is {x = width, y = height, z = channels,
 *     data}, row-major interleaved [y][x][z] float32 — the layout of the
 *     reference's vp_field (includes/fluid.hpp:17-22).  Velocity/pressure
 *     fields use channels 0 = u, 1 = v, 2 = pressure, 3 = divergence term.
 *   - Every `data` pointer passed to a pfs_* compute entry point is DEVICE
 *     memory (hipMalloc or equivalent).  includes/fluid.hpp's host-pointer
 *     flavour is provided by the C++ header include/fluid.hpp on top of this.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is enqueued asynchronously; nothing synchronises unless stated.
 *   - `iters` is the Jacobi sweep count, the reference's compile-time
 *     NUM_JACOBI_ITERS (includes/fluid.hpp:11, 30 there), here at run time.
 *   - Every function returns a status: PFS_OK (0) or a negative PFS_ERR_*;
 *     pfs_last_error() gives a message for the calling thread.  (The
 *     reference's functions return void and validate nothing.)
 *   - Like the reference, functions that ping-pong swap the `data` pointers
 *     inside the caller's pfs_field structs; the tables below say which.
 */
#ifndef PFS_H_
#define PFS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFS_OK 0
#define PFS_ERR_INVALID (-1)     /* bad argument: null pointer, size, channel count, aliasing */
#define PFS_ERR_HIP (-2)         /* a HIP runtime call failed */
#define PFS_ERR_NOMEM (-3)       /* device allocation failed */
#define PFS_ERR_UNSUPPORTED (-4) /* valid request outside what the library implements */

/* Same member layout as the reference's vp_field (includes/fluid.hpp:17-22). */
typedef struct pfs_field {
    int x;       /* width  */
    int y;       /* height */
    int z;       /* channels */
    float *data; /* device pointer, [y][x][z] */
} pfs_field;

/* ---- library ---------------------------------------------------------- */
const char *pfs_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *pfs_last_error(void);

/* ---- per-operator entry points ------------------------------------------
 * One per function of includes/fluid.hpp.  Each computes exactly what the
 * same-named function of src/fluid.cpp computes. */

/* includes/fluid.hpp:32 / src/fluid.cpp:24-70.  vp_out[0,1] <- bilinear
 * back-trace of vp[0,1] (periodic wrap); other channels of vp_out untouched.
 * Needs z >= 2; vp and vp_out must not alias. */
int pfs_advect(const pfs_field *vp, pfs_field *vp_out, float dt, void *stream);

/* includes/fluid.hpp:45 / src/fluid.cpp:72-127.  out[0..z) <- image advected
 * by vp's velocity sampled at the nearest cell; image and out same shape. */
int pfs_advect_color(const pfs_field *image, pfs_field *out, const pfs_field *vp, float dt,
                     void *stream);

/* includes/fluid.hpp:58 / src/fluid.cpp:129-196.  `iters` Jacobi smoothing
 * sweeps of channels 0,1 between vp and vp_out; swaps the two structs' data
 * pointers after every sweep but the last. */
int pfs_diffuse(pfs_field *vp, pfs_field *vp_out, float viscosity, float dt, int iters,
                void *stream);

/* includes/fluid.hpp:70 / src/fluid.cpp:198-208: a no-op in the reference. */
int pfs_add_forces(pfs_field *vp, const float *forces, void *stream);

/* Force injection, which the reference declares (includes/fluid.hpp:61-71:
 * "forces - A 3D array of force values for each pixel in the grid") but
 * leaves as an empty loop over every cell and channel (src/fluid.cpp:198-208)
 * whose call in the step is commented out (src/fluid.cpp:302, between
 * diffuse and computePressure).  Opt-in here:
 *   vp[j][i][k] += dt * forces[j][i][k]   for k < fz (fz <= vp->z),
 * forces a device array [vp->y][vp->x][fz]; the product is rounded, then the
 * sum (no fused multiply-add).  fz = 2 pushes the velocity only. */
int pfs_add_forces_apply(pfs_field *vp, const float *forces, int fz, float dt, void *stream);

/* includes/fluid.hpp:81 / src/fluid.cpp:210-267.  Divergence into channel 3
 * of both fields, then `iters` Jacobi sweeps of channel 2 with the same
 * pointer swaps as pfs_diffuse.  Needs z >= 4. */
int pfs_compute_pressure(pfs_field *vp, pfs_field *vp_out, float dt, int iters, void *stream);

/* includes/fluid.hpp:92 / src/fluid.cpp:269-296.  vp_out[0,1] <- vp[0,1]
 * minus the central-difference gradient of vp[2].  Needs z >= 3. */
int pfs_subtract_pressure_gradient(const pfs_field *vp, pfs_field *vp_out, float dt,
                                   void *stream);

/* ---- the step ----------------------------------------------------------- */

/* includes/fluid.hpp:109 / src/fluid.cpp:298-310.  One full step; on return
 * vp and tmp hold exactly what the reference's CPU step leaves in them (all
 * four channels of both), and the struct identities are restored as in the
 * reference.  Uses the fused gfx950 kernels when z == 4 and iters >= 1,
 * otherwise the per-operator sequence. */
int pfs_simulate_fluid_step(pfs_field *vp, pfs_field *tmp, float dt, float viscosity, int iters,
                            void *stream);

/* includes/fluid.hpp:118 / src/fluid.cpp:312-320.  Advect the dye, then swap
 * image->data and itmp->data. */
int pfs_advect_color_step(pfs_field *image, pfs_field *itmp, const pfs_field *vp, float dt,
                          void *stream);

/* One timestep of the reference driver (src/main.cpp:222-225, CPU branch
 * :236-239): simulate_fluid_step(vp, tmp) then advect_color_step(image,
 * itmp, vp), with the same results and the same image/itmp swap as the two
 * calls in sequence.  When the dye has the field's size and 4 channels, the
 * dye advection runs inside the step's last kernel (one pass fewer). */
int pfs_fluid_timestep(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                       float viscosity, int iters, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pfs_fluid_timestep, same arguments and results, with
 * the fused step's workspace in a device array the library holds per GPU
 * (PFS_NOALLOC_SCRATCH_BYTES, from the first call on a GPU until the process
 * ends) instead of a block it allocates.  Calls no hipMalloc or hipFree.  A
 * step whose workspace exceeds the array (16 B per field cell for
 * iters <= 8, 32 B beyond: up to 2M cells, or 1M past 8 sweeps) is rejected
 * with PFS_ERR_UNSUPPORTED before any launch.  Every call on one GPU uses
 * the same array: calls that overlap on one GPU (other streams or threads,
 * not ordered) race. */
#define PFS_NOALLOC_SCRATCH_BYTES (32u << 20)
int pfs_fluid_timestep_noalloc(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp,
                               float dt, float viscosity, int iters, void *stream);

/* includes/fluid.hpp:107.  Device-flavour step on raw device pointers (the
 * signature src/main.cpp:222 calls, plus iters and stream).  Same results
 * as pfs_simulate_fluid_step; *vp and *tmp are left un-swapped. */
int pfs_simulate_fluid_step_dev(float **vp, float **tmp, float dt, float viscosity, int vx,
                                int vy, int vz, int iters, void *stream);

/* includes/fluid.hpp:116.  Device-flavour dye step (src/main.cpp:225);
 * swaps *image and *itmp like the reference. */
int pfs_advect_color_step_dev(float **image, float **itmp, float **vp, float dt, int ix, int iy,
                              int iz, int vx, int vy, int vz, void *stream);

/* nsteps iterations of the reference driver loop (src/main.cpp:219-240) in
 * one call: pfs_fluid_timestep nsteps times, results identical.  On the
 * fused path runs of 16 and of 2 consecutive steps are captured once in
 * hipGraphs (kept, up to 16 per device, for later calls with the same
 * buffers, shapes, coefficients, tile side and stream workspace) and
 * replayed, which saves the per-launch host cost that bounds small fields.
 * image/itmp swap nsteps times, as the per-step calls would. */
int pfs_fluid_run(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                  float viscosity, int iters, int nsteps, void *stream);

/* Step-path selection for pfs_simulate_fluid_step / _dev (process-wide).
 *   PFS_PATH_AUTO     fused kernels when applicable (default)
 *   PFS_PATH_PER_OP   always the per-operator sequence (the reference's
 *                     structure, one launch per sweep) */
#define PFS_PATH_AUTO 0
#define PFS_PATH_PER_OP 1
int pfs_set_step_path(int path);
int pfs_get_step_path(void);

/* Tile side of the fused step's workgroups (process-wide).  No counterpart
 * in the reference, whose CUDA kernels use fixed 16x16 blocks
 * (src/fluid.cu:14-15).  Each workgroup of the fused kernels owns a side x
 * side tile of cells; results do not depend on the side.  Unless forced
 * here, a step takes the side pfs_autotune_step measured fastest for its
 * shape (field size, iters, dye fused or not), else the largest side whose
 * tiles fill every CU (a small field: 16).  side 0 clears a forced side. */
int pfs_set_step_tile(int side);
/* The side a step of this shape takes now (16, 32 or 64), or < 0 on error. */
int pfs_get_step_tile(int vx, int vy, int iters, int dye_fused);
/* Times pfs_fluid_timestep (image/itmp NULL: pfs_simulate_fluid_step) on
 * scratch copies of the buffers, `reps` steps per side after one untimed
 * step, for the sides 64, 32 and 16 in that order (ms[0..2], per step, if
 * ms is non-NULL with n_ms >= 3), and records the fastest for this shape:
 * later steps of the shape take it.  The caller's buffers are not changed.
 * Allocates the scratch copies; synchronises the stream. */
int pfs_autotune_step(const pfs_field *vp, const pfs_field *tmp, const pfs_field *image,
                      const pfs_field *itmp, float dt, float viscosity, int iters, int reps,
                      void *stream, int *side, float *ms, int n_ms);
/* Forgets every side pfs_autotune_step recorded. */
int pfs_clear_step_tuning(void);

/* Device scratch of the fused step: eight float planes of vx*vy cells (the
 * pressure start, the divergence, and two iterate planes each for diffusion
 * and pressure when iters exceeds one launch's sweeps).  The library keeps
 * one growable block per (device, stream); its size in bytes for a vx x vy
 * field is returned here. */
size_t pfs_step_workspace_bytes(int vx, int vy);

/* Per-kernel timing of one fused step (image/itmp NULL) or timestep
 * (pfs_fluid_timestep) on `stream`, measured with HIP events recorded on
 * that stream around each kernel group, the group launched `reps` times
 * back to back (each group is idempotent: it never reads what it writes):
 * ms[0] = advect+diffusion+divergence kernel(s), ms[1] =
 * pressure+projection(+dye) kernel(s), per launch of the group.  The
 * buffers advance by exactly one step.  Synchronises the stream. */
int pfs_profile_step(pfs_field *vp, pfs_field *tmp, pfs_field *image, pfs_field *itmp, float dt,
                     float viscosity, int iters, int reps, void *stream, float *ms, int n_ms);

/* ---- batched scenes --------------------------------------------------------
 * No counterpart in the reference, which steps one scene per call.  A batch
 * holds many independent scenes (caller-owned device buffers, as for
 * pfs_fluid_timestep) and advances all of them by one timestep each with one
 * launch per kernel stage; the launch sequences of 8 and of 2 steps are
 * replayed from hipGraphs while membership is unchanged (in them each step's
 * separate dye pass runs beside the next step's diffusion kernel; the batch
 * makes one more non-blocking stream of default priority for capturing that
 * branch).  Every scene's results are bit-identical to
 * stepping it alone with pfs_fluid_timestep.  Scenes need 4-channel,
 * 16-byte aligned velocity fields and 1 <= iters <= 512; the dye may have any
 * size and channel count. */
typedef struct pfs_batch pfs_batch;

/* stream NULL: the batch creates and owns a non-blocking stream. */
int pfs_batch_create(pfs_batch **out, void *stream);
/* Waits for the batch's work, frees its workspaces and tables. */
int pfs_batch_destroy(pfs_batch *b);
/* Adds a scene; the batch allocates its workspace (pfs_batch_workspace_bytes).
 * *slot identifies it until pfs_batch_remove. */
int pfs_batch_add(pfs_batch *b, const pfs_field *vp, const pfs_field *tmp, const pfs_field *image,
                  const pfs_field *itmp, float dt, float viscosity, int iters, int *slot);
/* Removes the scene from later steps without waiting: steps already queued
 * still read and write its buffers, so the caller keeps them until
 * pfs_batch_sync (the batch releases its workspace once those steps end). */
int pfs_batch_remove(pfs_batch *b, int slot);
/* Enqueues nsteps timesteps of every scene (simulate_fluid_step +
 * advect_color_step each); returns without waiting. */
int pfs_batch_step(pfs_batch *b, int nsteps);
int pfs_batch_sync(pfs_batch *b);
/* The scene's current dye buffer (the reference swaps image and itmp every
 * step), the other one, and the number of steps taken. */
int pfs_batch_scene(pfs_batch *b, int slot, pfs_field *image, pfs_field *itmp,
                    long long *steps);
int pfs_batch_count(pfs_batch *b);
size_t pfs_batch_workspace_bytes(pfs_batch *b);
/* 1 (default): replay steps from a hipGraph; 0: launch every kernel. */
int pfs_batch_set_graphs(pfs_batch *b, int enable);
void *pfs_batch_stream(pfs_batch *b);
/* Tile side of the batch's fused kernels (one side for every scene of a
 * launch; results do not depend on it).  side 0 (the default): the largest
 * side with a tile for every two CUs over all live scenes, else 16, chosen
 * again whenever membership changes; else the forced side. */
int pfs_batch_set_tile(pfs_batch *b, int side);
/* The side the next step uses (16, 32 or 64), or < 0 on error. */
int pfs_batch_tile(pfs_batch *b);
/* Times 2*reps steps of the batch for the sides 64, 32 and 16 (ms[0..2] per
 * step if ms is non-NULL with n_ms >= 3) and forces the fastest on the batch
 * (pfs_batch_set_tile(b, 0) returns to the automatic choice).  The scenes end
 * as they began: their buffers are snapshotted into a device allocation
 * of their size and copied back after each side.  Synchronises the batch. */
int pfs_batch_autotune(pfs_batch *b, int reps, int *side, float *ms, int n_ms);

/* ---- precision form --------------------------------------------------------
 * No counterpart in the reference, whose Jacobi loops run a fixed count in
 * FP32 (src/fluid.cpp:151,239).  Same step and buffer contract as
 * pfs_simulate_fluid_step with K = max_iters diffusion sweeps, except:
 * compensated (twice-FP32) stencil sums in both sweeps; an FP64 residual
 * ||b + pL + pR + pT + pB - 4p|| of every pressure iterate; and the pressure
 * iteration stops at the first iterate i >= 1 with residual <= tol * ||b||
 * (Kp <= max_iters sweeps), decided on the device.  With Kp = K the result
 * has the reference's layout: tmp = (E, p^K, b), vp = (E - grad p^K,
 * p^(K-1), b).  The FP64 sums are folded in a fixed order: the report, the
 * exit and the outputs are the same on every run.  vp and tmp must be
 * 16-byte aligned. */
#define PFS_PRECISE_MAX_ITERS 64
typedef struct pfs_precise_report {
    int iters_used;                                    /* Kp */
    double b_norm;                                     /* ||b||_2 */
    double residual_norm;                              /* ||r(p^Kp)||_2 */
    double residual_history[PFS_PRECISE_MAX_ITERS + 1]; /* ||r(p^s)||, s = 0.. */
} pfs_precise_report;

/* Device workspace bytes for a vx x vy field (fp64 = 1: for the FP64 twin). */
size_t pfs_precise_workspace_bytes(int vx, int vy, int fp64);
/* report: if non-NULL, filled after the stream is synchronised. */
int pfs_simulate_fluid_step_precise(pfs_field *vp, pfs_field *tmp, float dt, float viscosity,
                                    int max_iters, double tol, void *workspace,
                                    size_t workspace_bytes, void *stream,
                                    pfs_precise_report *report);
/* Schedule of the precision form's temporally blocked kernels,
 * process-wide: sweeps per launch (4 or 8; 0: the default, 8), tile side
 * (16, 32, 64; 0: the plain step's rule) and threads of the pressure
 * workgroups at tile side 64 (512, two per CU, or 1024; 0: the default,
 * 512).  The outputs do not depend on it; the FP64 residual sums are
 * partitioned by it (the same bits on every run of one schedule). */
int pfs_set_precise_schedule(int sweeps_per_launch, int tile, int pressure_threads);
/* FP64 twin: the same step and contract on double [vy][vx][4] device buffers
 * with plain arithmetic and given sweep counts (1 <= iters_pressure <=
 * iters_diffuse); with equal counts it is the reference's algorithm in
 * double.  The yardstick the precision form's error is measured against. */
int pfs_simulate_fluid_step_f64(double *vp, double *tmp, int vx, int vy, double dt,
                                double viscosity, int iters_diffuse, int iters_pressure,
                                void *workspace, size_t workspace_bytes, void *stream);

/* ---- particle layer ----------------------------------------------------------
 * No counterpart in the reference, which has no particles.  Tracers (SoA
 * float x[n], y[n] in field-cell units, periodic [0, W) x [0, H)) carried by
 * the grid's velocity, and a uniform neighbour grid over them: cell keys, a
 * stable radix sort, cell start/end tables, neighbour lists and an SPH
 * density.  All integer outputs are deterministic (a particle's neighbours
 * are listed by 3x3 cell rows, then columns, then sorted position). */
typedef struct pfs_ngrid {
    int n;                    /* particles */
    float domain_w, domain_h; /* periodic domain, same units as x, y */
    float radius;             /* neighbour radius r */
    int gx, gy;               /* cells per axis (>= 3, each at least r wide) */
    float kx, ky;             /* gx / domain_w, gy / domain_h: cell = (int)(x * kx) */
} pfs_ngrid;

/* One forward step: x += (dt * u(x, y)) / W, y += (dt * v(x, y)) / H, with u, v
 * bilinear in vp (channels 0, 1) at the position, corner keys and blend as
 * src/fluid.cpp:55-66; positions wrap into [0, W) x [0, H). */
int pfs_particles_advect(float *x, float *y, int n, const pfs_field *vp, float dt, void *stream);
/* Fills *g: gx = floor(W / r) (lowered until cells are >= r wide), likewise gy. */
int pfs_ngrid_init(pfs_ngrid *g, int n, float domain_w, float domain_h, float radius);
/* Workspaces of pfs_ngrid_* are 4-byte aligned, of this many bytes. */
size_t pfs_ngrid_workspace_bytes(const pfs_ngrid *g);
/* keys[n]: cell of particle i; order[n]: particle ids sorted by (key, id);
 * cell_start/cell_end[gx*gy]: [start, end) of each cell in sorted order
 * (an empty cell has start = end = its lower bound, so the cells of a row
 * are consecutive ranges); xs, ys[n]: positions in sorted order.  The
 * workspace keeps the sorted keys for pfs_ngrid_neighbors / _density. */
int pfs_ngrid_build(const pfs_ngrid *g, const float *x, const float *y, uint32_t *keys,
                    uint32_t *order, uint32_t *cell_start, uint32_t *cell_end, float *xs, float *ys,
                    void *workspace, size_t workspace_bytes, void *stream);
/* counts[i]: neighbours of particle i within r (periodic distance, d^2 < r^2,
 * itself excluded); offsets[i]: their exclusive scan; *total (device): the
 * sum.  With nbr non-NULL also the lists, nbr[offsets[i] + k] (particle
 * ids); this waits for *total to check nbr_capacity.  With order NULL,
 * particles are identified by sorted position instead of id (for callers
 * that keep their particles in cell order, pfs_gather_u32). */
int pfs_ngrid_neighbors(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                        const uint32_t *cell_start, const uint32_t *cell_end, uint32_t *counts,
                        uint32_t *offsets, uint32_t *nbr, size_t nbr_capacity, uint32_t *total,
                        void *workspace, size_t workspace_bytes, void *stream);
/* pfs_ngrid_neighbors' offsets, *total and lists from counts the caller
 * already holds for the same positions, order and grid (those of
 * pfs_ngrid_density with counts, say), without counting again.  Counts that
 * are not the traversal's give wrong lists, never writes past nbr_capacity. */
int pfs_ngrid_lists(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                    const uint32_t *cell_start, const uint32_t *cell_end, const uint32_t *counts,
                    uint32_t *offsets, uint32_t *nbr, size_t nbr_capacity, uint32_t *total,
                    void *workspace, size_t workspace_bytes, void *stream);
/* rho[i] = (mass * 4/(pi r^8)) * S, S the FP32 sum over i and its neighbours
 * of (r^2 - d^2)^3 (2D poly6) in the traversal order (i where its cell's
 * range meets it); counts (may be NULL) gets the neighbour counts of the same
 * traversal. */
int pfs_ngrid_density(const pfs_ngrid *g, const float *xs, const float *ys, const uint32_t *order,
                      const uint32_t *cell_start, const uint32_t *cell_end, float mass, float *rho,
                      uint32_t *counts, void *workspace, size_t workspace_bytes, void *stream);
/* Position-based fluids (Macklin & Mueller 2013): Jacobi iterations of the
 * 2D density constraint C_i = rho_i / rho0 - 1 on the sorted positions
 * (xs, ys, in place), with kernel radius h = r, the cell tables of the last
 * pfs_ngrid_build (whose workspace ngrid_ws holds the sorted keys) and the
 * neighbour lists of those positions (pfs_ngrid_neighbors with order NULL:
 * offsets, counts and lists by sorted position), evaluated at the current
 * iterate:
 *   rho_i = mass * 4/(pi h^8) * ((h^2)^3 + sum_{j in list i} (h^2 - r_ij^2)^3)
 *   lambda_i = -C_i / (sum_k |grad_k C_i|^2 + eps)
 *   x_i += (1/rho0) * sum_{j in list i} (lambda_i + lambda_j) grad W_spiky(x_i - x_j),
 *   grad W_spiky(d) = -30/(pi h^5) (h - |d|)^2 d / |d|  for 0 < |d| < h, else 0,
 * every update of an iteration from the same iterate; positions wrap.  Sums
 * run in list order, FP32.  residuals[k] (device, max_iters + 1 doubles) =
 * sum_i C_i^2 of iterate k in FP64 (-1 for iterates not reached); the solve
 * stops at the first iterate with sum C^2 <= tol^2 * n (tested on the
 * device).  lambda (device, n) gets the last iterate's lambdas; *iters_used
 * (device, may be NULL) the number of updates applied.  ws:
 * pfs_pbf_workspace_bytes, 256-byte aligned. */
#define PFS_PBF_MAX_ITERS 64
size_t pfs_pbf_workspace_bytes(const pfs_ngrid *g, int max_iters);
int pfs_pbf_solve(const pfs_ngrid *g, float *xs, float *ys, const uint32_t *cell_start,
                  const uint32_t *cell_end, const uint32_t *nbr_offsets, const uint32_t *nbr_counts,
                  const uint32_t *nbr, float mass, float rho0, float eps, int max_iters, double tol,
                  float *lambda, double *residuals, int *iters_used, void *ngrid_ws,
                  size_t ngrid_ws_bytes, void *ws, size_t ws_bytes, void *stream);
/* out[p] = in[index[p]]: permutes per-particle data into sorted order. */
int pfs_gather_u32(const uint32_t *index, const uint32_t *in, uint32_t *out, int n, void *stream);
/* Exclusive prefix sum of n uint32 (wavefront DPP scans); *total (device,
 * may be NULL) gets the sum.  Workspace: (ceil(n / 4096) + 1) * 4 bytes. */
int pfs_scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total,
                           void *workspace, size_t workspace_bytes, void *stream);

/* ---- PNG files of the reference's harness ---------------------------------
 * The reference reads its input image and velocity field with libpng's
 * simplified API as 8-bit RGBA and converts each byte to float (byte/255.0),
 * and writes every frame as 8-bit RGBA with (png_byte)(x*255.0)
 * (includes/utils.hpp:32-104, 115-150; src/main.cpp:126-140, 59-67).  These
 * entry points do the same on the host without libpng: any PNG colour type,
 * bit depth and interlace, converted to RGBA8 as libpng 1.6's simplified
 * API does (gamma to sRGB included); floats are [h][w][4]. */
int pfs_png_size(const char *path, int *width, int *height);
/* out: width*height*4 bytes. */
int pfs_png_read_rgba8(const char *path, uint8_t *out, size_t n_bytes);
/* out: width*height*4 floats, each (float)((double)byte / 255.0) (utils.hpp:83-85). */
int pfs_png_read(const char *path, float *out, size_t n_floats);
int pfs_png_write_rgba8(const char *path, const uint8_t *rgba, int width, int height);
/* data: width*height*4 floats; byte = (png_byte)(x * 255.0) as the reference
 * evaluates it on x86-64 (utils.hpp:129): truncation, low byte of the 32-bit
 * integer for values outside [0, 256). */
int pfs_png_write(const char *path, const float *data, int width, int height);

/* ---- the simulation driver -------------------------------------------------
 * src/main.cpp:71-280 on the GPU: argv as the reference's
 *   <n_timesteps> <delta_t> <viscosity> <input_image> <velocity_field> [output_dir]
 * plus options (--iters K, --every N, --npy, --dump DIR, --per-op,
 * --buffers N, --quiet); the same messages; frames copied out on a second
 * stream through pinned buffers while the steps run on.  Returns the
 * process exit status.  The pfs_sim executable is main() around it. */
int pfs_sim_main(int argc, char **argv);

/* ---- debug: the integer "neighbour lists" ------------------------------
 * Back-trace cell keys of advect (src/fluid.cpp:55-57): keys[(j*w+i)*4 +
 * {0,1,2,3}] = {i0, j0, i1, j1} for every cell.  `keys` is a device int32
 * array of 4*w*h. */
int pfs_debug_advect_keys(const pfs_field *vp, float dt, int32_t *keys, void *stream);
/* Keys of advect_color (src/fluid.cpp:89-90,112-113): keys[(j*iw+i)*6 +
 * {0..5}] = {vi, vj, i0, j0, i1, j1} per pixel; device int32 array of 6*iw*ih. */
int pfs_debug_advect_color_keys(const pfs_field *image, const pfs_field *vp, float dt,
                                int32_t *keys, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PFS_H_ */
