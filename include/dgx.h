/*
 * dgx.h -- C ABI of libdgx, the MI355X (gfx950) DG residual and explicit
 * time-step library.
 *
 * The reference (DGPR) has no plugin API: its GPU path is a set of free
 * functions and POD structs #included into solver/cudg.cu.  Each entry point
 * below names the reference interface it replaces (file:line in DGPR).
 * All functions return a dgx_status; none aborts the process.  Every buffer
 * is a plain pointer; sizes are implied by the mesh (nelem*5*(p+1)^3 doubles
 * for a state) or passed explicitly.
 *
 * State layout at this boundary is the reference's simstate layout
 * U[e][r][bf] (geometry_common.cpp:467-470): conserved variable r of
 * equispaced-Lagrange coefficient bf on element e, elements ordered
 * e = nx*ny*k + nx*j + i, coefficients ordered z-y-x.
 */
#ifndef DGX_H
#define DGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dgx_status
{
  DGX_OK              = 0,
  DGX_ERR_INVALID     = 1, /* bad argument (null pointer, size, order)     */
  DGX_ERR_UNSUPPORTED = 2, /* valid input outside what the kernels cover   */
  DGX_ERR_HIP         = 3, /* a HIP runtime call failed                    */
  DGX_ERR_NAN         = 4, /* residual norm is NaN ("went unstable")       */
  DGX_ERR_IO          = 5, /* .dg file could not be read or written        */
  DGX_ERR_NODEVICE    = 6  /* no HIP device visible                        */
} dgx_status;

/* btype (boundaries.cpp:26-35): negative codes stored in bounds[6] */
enum
{
  DGX_BT_INTERIOR         = -1,
  DGX_BT_PERIODIC         = -2,
  DGX_BT_FREESTREAM       = -3,
  DGX_BT_INVISCID_WALL    = -4,
  DGX_BT_SUBSONIC_INFLOW  = -5,
  DGX_BT_SUBSONIC_OUTFLOW = -6,
  DGX_BT_NO_SLIP_WALL     = -7
};

/* time_scheme (time.cpp:23-27) */
enum
{
  DGX_RK4    = 0,
  DGX_TVDRK3 = 1
};

/* parameters (solver_params.cpp:22-36), passed by value to the kernels
 * exactly as the reference passes it to cuda_residual. */
typedef struct dgx_params
{
  double gamma;     /* ratio of specific heats        */
  double R;         /* specific gas constant           */
  double mu;        /* dynamic viscosity               */
  double c;         /* reference speed of sound        */
  double T;         /* reference static temperature    */
  double P;         /* reference static pressure       */
  double Ufr[5];    /* free-stream conserved state     */
  double source[5]; /* constant source term            */
  double eta;       /* BR2 stabilisation parameter     */
} dgx_params;

/* The persistent part of core_geometry (geometry_common.cpp:25-83), i.e.
 * what a .dg file stores (io.cpp:33-89): orders, element counts, boundary
 * codes per domain side (-x +x -y +y -z +z) and element nodes
 * [e][(q+1)^3][3].  Face lists are implied by the structured layout exactly
 * as shared_geometry::precompute_face_geometry derives them
 * (geometry_shared.cpp:156-267). */
typedef struct dgx_mesh
{
  uint64_t p, q;
  uint64_t nx, ny, nz;
  int64_t bounds[6];
  const double* nodes;
} dgx_mesh;

typedef struct dgx_ctx dgx_ctx;

/* ---- context: device geometry + residual workspace ----------------------
 * Replaces cuda_device_geometry::cuda_device_geometry(shared_geometry&)
 * (solve_cuda.cu:171) together with cuda_residual_workspace(...)
 * (solve_cuda.cu:303): uploads the mesh, allocates the state and the stage
 * buffers of both RK schemes on the current HIP device and creates one
 * stream.  Supported: every order the reference accepts, 0 <= p <= 8 and
 * 1 <= q <= 8 (MAX_ORDER, constants.cpp:21, geometry_distributed.cpp:
 * 819-830).  1 <= p, q <= 3 run the tuned kernels; a q = 1 mesh there whose
 * elements are all parallelepipeds runs the constant-Jacobian kernels
 * (chosen here, nothing to set).  The other orders run the general-order
 * kernels (stage tile 2, one Gauss node per lane); the adjoint covers
 * 1 <= p, q <= 3 and returns DGX_ERR_UNSUPPORTED outside. */
dgx_status dgx_create(const dgx_mesh* mesh, const dgx_params* params,
                      dgx_ctx** out);

/* Replaces free_cuda_device_geometry (solve_cuda.cu:267) and
 * free_cuda_residual_workspace (solve_cuda.cu:323). */
dgx_status dgx_destroy(dgx_ctx* ctx);

/* Replaces the parameters argument of cuda_residual / RK4_cuda
 * (solve_cuda.cu:1235, time_cuda.cu:152): the reference state can change
 * between calls (e.g. viscosity sweeps). */
dgx_status dgx_set_params(dgx_ctx* ctx, const dgx_params* params);

/* Replaces shuffle_state + cudaMemcpy H2D (cudg.cu:97-106): U_host is a
 * simstate [e][5][nbf]. */
dgx_status dgx_set_state(dgx_ctx* ctx, const double* U_host);

/* Replaces cudaMemcpy D2H + unshuffle_state (cudg.cu:131-136). */
dgx_status dgx_get_state(dgx_ctx* ctx, double* U_host);

/* Device-pointer forms of the two transfers above: U_dev is a device
 * buffer in simstate layout. */
dgx_status dgx_set_state_device(dgx_ctx* ctx, const double* U_dev);
dgx_status dgx_get_state_device(dgx_ctx* ctx, double* U_dev);

/* Replaces cuda_residual(state, geom, params, wsp, residual, f)
 * (solve_cuda.cu:1235-1282) with the CPU semantics of residual()
 * (solve.cpp:150-807): R = volume - source + faces, f = -Minv R.  All three
 * are device buffers in simstate layout; R_dev may be NULL.  Runs on the
 * context's stream and returns after completion. */
dgx_status dgx_residual_device(dgx_ctx* ctx, const double* U_dev,
                               double* R_dev, double* f_dev);

/* Host-buffer convenience form: evaluates the residual of the context's
 * current state; R_host may be NULL. */
dgx_status dgx_residual(dgx_ctx* ctx, double* R_host, double* f_host);

/* Replaces the step function pointer of cudg.cu:75-91, i.e. RK4_cuda
 * (time_cuda.cu:152) and tvdRK3_cuda (time_cuda.cu:126), with the CPU
 * stage algebra of RK4_distributed / tvdRK3_distributed
 * (time_distributed.cpp:46-102).  Advances the context's state by nsteps
 * steps of size dt.  If norms_host is not NULL it receives, per step, the
 * RMS of R at the first stage (residual_norm, time_distributed.cpp:23-44)
 * and a NaN norm stops stepping with DGX_ERR_NAN. */
dgx_status dgx_step(dgx_ctx* ctx, int scheme, double dt, uint64_t nsteps,
                    double* norms_host);

/* Enqueue-only form of dgx_step without norms: launches the stages of
 * nsteps steps on the context's stream and returns without waiting. */
dgx_status dgx_step_async(dgx_ctx* ctx, int scheme, double dt,
                          uint64_t nsteps);

/* Enqueue-only form of dgx_step with norms: the RMS of R at the first stage
 * of step s (residual_norm, time_distributed.cpp:23-44, called by
 * RK4_distributed / tvdRK3_distributed every step) goes to norms_dev[s], a
 * device buffer of nsteps doubles, without a host round trip per step.  The
 * caller reads them after dgx_sync and applies the reference's NaN check
 * itself (dgx_step stops at the first NaN). */
dgx_status dgx_step_norms_async(dgx_ctx* ctx, int scheme, double dt, uint64_t nsteps,
                                double* norms_dev);

/* Enqueues a device-to-device copy of bytes (a multiple of 16) on the
 * context's stream with a plain streaming kernel.  Replaces the reference's
 * state copies between device buffers (cudaMemcpy device-to-device is not
 * used by the reference, which copies through the host, cudg.cu:115-136);
 * it also serves as the known-size kernel for memory-traffic calibration. */
dgx_status dgx_device_copy(dgx_ctx* ctx, void* dst, const void* src,
                           uint64_t bytes);

/* Waits for all work on the context's stream. */
dgx_status dgx_sync(dgx_ctx* ctx);

/* The context's hipStream_t (as void*), for callers that time or order
 * work around the library with HIP events. */
dgx_status dgx_get_stream(dgx_ctx* ctx, void** stream);

/* Sizes: out[0]=nelem, out[1]=nbf=(p+1)^3, out[2]=state length
 * nelem*5*nbf, out[3]=device bytes held by the context. */
dgx_status dgx_get_sizes(const dgx_ctx* ctx, uint64_t out[4]);

/* Human-readable text of the last error recorded on ctx (or of the last
 * failed dgx_create when ctx is NULL). */
const char* dgx_last_error(const dgx_ctx* ctx);

/* ---- derived outputs ----------------------------------------------------- */

/* output_type (output.cpp:23-30) */
enum
{
  DGX_OUT_MACH = 0,
  DGX_OUT_RHO  = 1,
  DGX_OUT_U    = 2,
  DGX_OUT_V    = 3,
  DGX_OUT_W    = 4
};

/* Replaces eval_output (render/vrtxgen.cpp:25-59) over the sub-grid of
 * compute_metadata_cpu (vrtxgen.cpp:294-330): the output var of the
 * context's state at the (N+1)^3 equispaced reference points (i/N, j/N, k/N)
 * of every element, out[e][(N+1)^2 k + (N+1) j + i].  The Mach number keeps
 * the reference's pressure (gamma - 1)(rho E - |u|^2 / 2).  1 <= N <= 64.
 * out_dev is a device buffer of nelem*(N+1)^3 doubles; the host form copies
 * the result back. */
dgx_status dgx_sample_output_device(dgx_ctx* ctx, int var, int N, double* out_dev);
dgx_status dgx_sample_output(dgx_ctx* ctx, int var, int N, double* out_host);

/* Replaces update_mean(U, mean, time_step, mean_start) of the driver loop
 * (output.cpp:39-49, dg.cpp:123-131) on the context's current state: the
 * mean is kept on the device and restarts (zeroed, then updated) when
 * time_step == mean_start.  time_step < mean_start is DGX_ERR_INVALID. */
dgx_status dgx_update_mean(dgx_ctx* ctx, uint64_t time_step, uint64_t mean_start);

/* The running mean in simstate layout (DGX_ERR_INVALID before the first
 * dgx_update_mean or dgx_set_mean). */
dgx_status dgx_get_mean(dgx_ctx* ctx, double* mean_host);

/* Sets the running mean from a simstate, e.g. the "mean" output of a .dg
 * file that a restarted run continues to average into (the CPU driver keeps
 * the mean it read, dg.cpp:44-72, 122-131). */
dgx_status dgx_set_mean(dgx_ctx* ctx, const double* mean_host);

/* ---- batched ensembles (no counterpart in the reference, which runs one
 * case per process, dg.cpp / cudg.cu) ----------------------------------------
 * An ensemble owns one context per member (meshes of any size, p and q; own
 * parameters).  dgx_ensemble_step advances every member by nsteps steps of
 * its own dt.  Per RK stage, by default (layout 1), all q = 1 members of
 * every order run as ONE launch of the ensemble stage kernel over the
 * concatenation of their work units (rows or tiles of one member's elements,
 * largest members first; affine and curved members form two such launches;
 * q > 1 members run one batched launch per (p, q)); with
 * dgx_ensemble_set_layout(ens, 2) that kernel runs once per (order, affine)
 * class of the q = 1 members instead, and with dgx_ensemble_set_layout(ens,
 * 0) every (p, q, affine) group runs its own batched launch of the one-row
 * kernel.  Members of an order outside 1 <= p, q <= 3 run one launch of the
 * general-order kernel each, in every layout.  With use_graph != 0 one step of the whole ensemble is captured
 * once in a hipGraph and replayed.  Results equal those of dgx_step on each
 * member (the same arithmetic, to rounding) in every
 * layout.  The call only enqueues work on the ensemble's stream: call dgx_ensemble_sync before
 * using a member context again. */
typedef struct dgx_ensemble dgx_ensemble;
dgx_status dgx_ensemble_create(int ncase, const dgx_mesh* meshes, const dgx_params* params,
                               dgx_ensemble** out);
dgx_status dgx_ensemble_destroy(dgx_ensemble* ens);
/* The context of member i (set/get its state, residual, single-case steps). */
dgx_status dgx_ensemble_member(dgx_ensemble* ens, int i, dgx_ctx** ctx);
dgx_status dgx_ensemble_step(dgx_ensemble* ens, int scheme, const double* dt, uint64_t nsteps,
                             int use_graph);
/* Waits for the ensemble's stream; DGX_ERR_INVALID if a member met a
 * subsonic inflow without a Mach root. */
dgx_status dgx_ensemble_sync(dgx_ensemble* ens);
dgx_status dgx_ensemble_get_stream(dgx_ensemble* ens, void** stream);
/* For each layout (1, 2, 0) times every launch group at each occupancy
 * target and, for the ensemble stage kernel,
 * each tile (one-row units: 64/(p+1)^3 elements per workgroup, one node per
 * lane; tiled units: 2, 6 or 8 elements per workgroup at p = 3, 2, 1, up to
 * 3 nodes per lane), with reps launches of the first stage (which writes
 * stage buffers only), keeps the fastest per group, then keeps the layout
 * whose first stage over all members is fastest (a later
 * dgx_ensemble_set_layout overrides it).  Writes each member's
 * occupancy target to waves_out[ncase] (may be NULL;
 * dgx_ensemble_member_tuning reads the full choice). */
dgx_status dgx_ensemble_autotune(dgx_ensemble* ens, int scheme, const double* dt, int reps,
                                 int* waves_out);
/* Layout of the q = 1 members' launches: 1 (the default) one launch per
 * stage for all orders; 2 one launch per order; 0 one batched launch per
 * (p, q, affine) group. */
dgx_status dgx_ensemble_set_layout(dgx_ensemble* ens, int layout);
/* The launch member i runs in: layout (0 for a batched launch, else the
 * layout 1 or 2 of dgx_ensemble_set_layout), tile (0/1; 2 for the
 * general-order kernel) and occupancy target (waves per SIMD). */
dgx_status dgx_ensemble_member_tuning(const dgx_ensemble* ens, int i, int* layout, int* tile,
                                      int* waves);
const char* dgx_ensemble_last_error(const dgx_ensemble* ens);

/* ---- discrete adjoint (no counterpart in the reference) --------------------
 * Exact transposes of the device's discrete residual and RK4 stages, on
 * meshes with any boundary types (walls, inflow, outflow, free stream: the
 * boundary states are differentiated with the face physics). */

/* J = volume integral of the kinetic energy |m|^2 / (2 rho) of the context's
 * state, with the (p+1)^3 Gauss rule of the mass matrix. */
dgx_status dgx_kinetic_energy(dgx_ctx* ctx, double* J);

/* g = (df/dU)^T w at the context's state: the vector-Jacobian product of
 * residual()'s f = -Minv R (solve.cpp:150-807); w and g in simstate layout. */
dgx_status dgx_residual_vjp(dgx_ctx* ctx, const double* w_host, double* g_host);

/* Allocation-free form of dgx_residual_vjp on device buffers (no
 * counterpart in the reference, which has no adjoint): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * optimisation loops, harnesses that forbid hipMalloc).  w_dev and g_dev are
 * device buffers in simstate layout; the workspace (3 states + 2 trace-record
 * sets) comes from a fixed 64 MiB device array that the library holds on
 * every device it runs on, so meshes needing more are rejected with
 * DGX_ERR_INVALID (at p = 3 about 3300 elements fit, at p = 2 about 6600).
 * Calls on one device share that array: calls that overlap on one device
 * (other threads, other contexts) race and must be serialised by the
 * caller.  Returns after completion; results equal dgx_residual_vjp's. */
dgx_status dgx_residual_vjp_device_noalloc(dgx_ctx* ctx, const double* w_dev, double* g_dev);

/* Runs nsteps classic RK4 steps of size dt from the context's state (which
 * ends at the final state, as after dgx_step) and returns the final kinetic
 * energy J, its gradient with respect to the initial state (simstate layout,
 * may be NULL) and dJ/dmu (may be NULL), by the discrete adjoint.
 * interval >= 1: a checkpoint every `interval` steps; in the reverse sweep
 * each segment is replayed once into a store of `interval` steps' stage
 * states with their trace records (and f, for dJ/dmu; the last segment is
 * written there by the forward sweep itself), or, when that store
 * does not fit, its step-start states are recomputed and each step's stages
 * rebuilt.  interval = 0: every stage state with its records is kept from
 * the forward sweep when nsteps * 4 * (state + records [+ state]) fits in
 * three quarters of the free device memory, so nothing is recomputed;
 * otherwise checkpoints every ceil(sqrt(nsteps)) steps as above.  All
 * schedules give the same gradients bit for bit.  The store stays allocated
 * on the context for later calls (a buffer that is large enough is reused;
 * freeing tens of GB can take seconds) until dgx_adjoint_release or
 * dgx_destroy; the free-memory test counts the held buffer as free, and a
 * call that checkpoints frees a held buffer too small for its segments
 * before its own allocations. */
dgx_status dgx_adjoint_rk4(dgx_ctx* ctx, double dt, uint64_t nsteps, uint64_t interval,
                           double* J, double* dJdU0_host, double* dJdmu);

/* Frees the stored-stage buffer that dgx_adjoint_rk4 keeps on the context
 * (no counterpart in the reference, which has no adjoint).  No-op when none
 * is held. */
dgx_status dgx_adjoint_release(dgx_ctx* ctx);
/* Caps the stored-stage buffer of dgx_adjoint_rk4 at `bytes` (0: no cap
 * beyond three quarters of the free device memory; no counterpart in the
 * reference).  A schedule whose store would exceed the cap falls to the next
 * one (every stage stored -> one replay per checkpoint segment -> stages
 * rebuilt per step); a held buffer above the cap is freed.  Gradients are
 * the same bit for bit on every schedule. */
dgx_status dgx_adjoint_set_store_limit(dgx_ctx* ctx, uint64_t bytes);

/* ---- precision form (no counterpart in the reference: its CPU solver is
 * FP64 and its CUDA driver FP32 throughout, cudg.cu:26-27) ----------------- */

typedef enum dgx_precision
{
  DGX_PREC_FP64       = 0, /* everything in FP64 (default)                  */
  DGX_PREC_MIXED      = 1, /* FP32 trace records; p = 3 tiled kernel: volume
                              contractions in FP32 on the matrix cores       */
  DGX_PREC_MIXED_BF16 = 2, /* as MIXED, the contractions as split-BF16
                              (3-term) products on the BF16 matrix cores    */
  DGX_PREC_MIXED_RECORDS = 3 /* FP32 trace records only; every contraction
                              in FP64                                       */
} dgx_precision;

/* The mixed forms store the face trace records that cross element faces
 * between stages (the interpolated trace state and the gradient traction,
 * computed in FP64 and rounded once) in FP32, and in the tiled p = 3 kernel
 * run the volume contractions (the reference-space gradient of density and
 * momentum, and the projection of the contravariant volume flux onto the
 * nodes) on the matrix cores with FP32 accumulation: FP32 operands
 * (DGX_PREC_MIXED) or each FP32 operand split into three BF16 terms, the six
 * leading products accumulated (DGX_PREC_MIXED_BF16).  DGX_PREC_MIXED_RECORDS
 * keeps the FP32 records and runs every contraction in FP64 (at p = 3 the
 * fastest and most accurate of the three: the contractions are a small part
 * of the kernel and the matrix-core forms add conversions).  Geometry, the Roe
 * flux and its eigen-structure, pressure/energy recovery, the viscous flux
 * physics, the inverse mass and the RK accumulation stay FP64.  Changing the
 * form drops the stored records (the next stage recomputes them).  The
 * adjoint entry points and ensemble steps need DGX_PREC_FP64
 * (DGX_ERR_UNSUPPORTED otherwise). */
dgx_status dgx_set_precision(dgx_ctx* ctx, int precision);
dgx_status dgx_get_precision(const dgx_ctx* ctx, int* precision);

/* ---- launch tuning (no counterpart in the reference, whose kernels have
 * fixed launch shapes, solve_cuda.cu:1235-1282) ---------------------------- */

/* The stage kernel is compiled for occupancy targets of 1, 2 and 3 waves per
 * SIMD (register budgets of 512, 256 and 168 VGPRs).  dgx_create picks a
 * default per (p, q); these set or read the target used by later stages. */
dgx_status dgx_set_tuning(dgx_ctx* ctx, int stage_waves);
dgx_status dgx_get_tuning(const dgx_ctx* ctx, int* stage_waves);

/* Times every compiled occupancy target of the current stage tile on this
 * context's mesh and state (reps launches each into scratch buffers; the
 * state is unchanged), keeps the fastest and returns it in *chosen (may be
 * NULL).  us_out (may be NULL) receives the average launch times in
 * microseconds at 1, 2 and 3 waves per SIMD (0 for a target the tile does
 * not compile). */
dgx_status dgx_autotune(dgx_ctx* ctx, int reps, int* chosen, double us_out[3]);

/* Stage tile: 0 runs the stage kernel that maps one row of whole elements
 * (one element at p = 3, 2 at p = 2, 8 at p = 1) to a wavefront and mixes
 * the faces of all directions in each face pass; 1 runs the tiled kernel:
 * several rows per wavefront (2 at p = 3, 3 at p = 2, 1 at p = 1), each lane
 * owning one node per row, faces processed one reference direction at a
 * time in whole-face subpasses (compiled for q = 1 meshes at 2 and 3 waves
 * per SIMD; selecting it moves a target of 1 to 2; DGX_ERR_UNSUPPORTED for
 * q > 1); 2 runs the general-order kernel (one element, or several at
 * p <= 2, per workgroup, one Gauss node per lane; every order, and the only tile outside 1 <= p, q <= 3,
 * where 0 and 1 return DGX_ERR_UNSUPPORTED; occupancy targets do not apply).
 * Results agree to rounding. */
dgx_status dgx_set_tile(dgx_ctx* ctx, int tile);
dgx_status dgx_get_tile(const dgx_ctx* ctx, int* tile);

/* Times every (tile, occupancy target) pair on this context's mesh as
 * dgx_autotune does and keeps the fastest; us_out[3 * tile + waves - 1]
 * (may be NULL) receives the times (0 where not compiled). */
dgx_status dgx_autotune_tiles(dgx_ctx* ctx, int reps, int* tile_out, int* waves_out,
                              double us_out[6]);

/* ---- host-side helpers (no device work) --------------------------------- */

/* make_reference_state (solver_params.cpp:40-70): P = c^2 rho/gamma,
 * T = P/(rho R), Ufr from (rho, u, v, w, P).  source may be NULL. */
dgx_status dgx_make_reference_state(double gamma, double R, double mu,
                                    double c, double rho, double u, double v,
                                    double w, double eta,
                                    const double* source, dgx_params* out);

/* generate_bump_nodes as called by bump_shared (geometry_common.cpp:494-576,
 * geometry_shared.cpp:275-292): nodes [e][(q+1)^3][3] of an nx*ny*nz
 * element box of size lx*ly*lz with wall-normal tanh stretching a and a
 * Gaussian bump on the lower y wall.  nodes_out holds
 * nx*ny*nz*(q+1)^3*3 doubles. */
dgx_status dgx_bump_nodes(uint64_t q, uint64_t nx, uint64_t ny, uint64_t nz,
                          double lx, double ly, double lz, double a,
                          double bmp_height, double bmp_center,
                          double bmp_variance, double* nodes_out);

/* Whether dgx_create will run the constant-Jacobian kernels on this mesh
 * (no counterpart in the reference, which stores per-point metrics for every
 * element, geometry_common.cpp:292-395): *affine = 1 when q = 1 and every
 * element is a parallelepiped, i.e. its bilinear and trilinear node terms
 * are below 1e-14 of its longest edge-vector component plus 8 ulps of its
 * largest coordinate. */
dgx_status dgx_mesh_is_affine(const dgx_mesh* mesh, int* affine);

/* Box nodes on [x0,x0+lx]x[y0,y0+ly]x[z0,z0+lz] with each node displaced by
 * a smooth field of relative amplitude amp (fraction of the local element
 * size) that is periodic in x and z and vanishes on the y=y0 and y=y0+ly
 * planes.  amp = 0 gives an affine box.  seed selects the field's phases. */
dgx_status dgx_box_nodes(uint64_t q, uint64_t nx, uint64_t ny, uint64_t nz,
                         double x0, double y0, double z0, double lx,
                         double ly, double lz, double amp, uint64_t seed,
                         double* nodes_out);

/* Taylor-Green-type initial state sampled at the Lagrange nodes (as
 * solver_init.cpp:43-76 samples at nodes):
 *   u = V0 sin x cos y cos z, v = -V0 cos x sin y cos z, w = 0,
 *   p = P0 + (rho0 V0^2/16)(cos 2x + cos 2y)(cos 2z + 2),
 *   rho = rho0 p/P0 (isothermal), plus uniform noise in [-noise, noise] on
 *   each velocity component from a splitmix64 stream seeded by seed.
 * U_out is a simstate [e][5][nbf]. */
dgx_status dgx_tgv_state(const dgx_mesh* mesh, double gamma, double rho0,
                         double V0, double P0, double noise, uint64_t seed,
                         double* U_out);

/* Channel initial state: laminar Poiseuille u = Uc(1 - y^2) on y in
 * [-1, 1], rho = rho0, p = P0, plus divergence-free noise built as the
 * curl of a random vector potential with Fourier modes |m| <= kmax
 * (periodic in x and z, damped to zero at the walls), scaled to an RMS of
 * frac*Uc over the nodes. */
dgx_status dgx_channel_state(const dgx_mesh* mesh, double gamma, double rho0,
                             double Uc, double P0, double frac, int kmax,
                             double lx, double lz, uint64_t seed,
                             double* U_out);

/* turbulent (solver_init.cpp:30-73), the "-turb" initial state of the
 * driver: free-stream density and energy; velocity component i is the
 * free-stream value plus amp * sum over L = l/2, l/4, l/8, l/16 of
 * prod_j sin(2 pi (x_j - 0.3 i)/L), amp = |u_fr|/10, at the Lagrange nodes.
 * U_out is a simstate [e][5][nbf]. */
dgx_status dgx_turbulent_state(const dgx_mesh* mesh, const dgx_params* params,
                               double l, double* U_out);

/* .dg checkpoint file (io.cpp:33-159): write_state with one output named
 * "state", and read_state of that output (of the first output so named).  dgx_read_dg_header fills
 * tstep and a mesh header (nodes NULL); dgx_read_dg then reads nodes and
 * state into caller buffers sized from that header. */
dgx_status dgx_write_dg(const char* path, uint64_t tstep, const dgx_mesh* mesh,
                        const double* U);
dgx_status dgx_read_dg_header(const char* path, uint64_t* tstep,
                              dgx_mesh* mesh);
dgx_status dgx_read_dg(const char* path, double* nodes_out, double* U_out);

/* write_state with several named outputs (io.cpp:33-89), e.g. "state" and
 * the running "mean" of the CPU driver (dg.cpp:122-131): keys[o] names the
 * simstate states[o]; keys are unique, 1..1023 characters.  Outputs are
 * written in the order the reference's map<simstate> visits them (djb2
 * bucket of the key, data_structures.cpp:537-673), so the file is byte for
 * byte the reference's ("state" before "mean"). */
dgx_status dgx_write_dg_outputs(const char* path, uint64_t tstep, const dgx_mesh* mesh,
                                uint64_t nout, const char* const* keys,
                                const double* const* states);

/* read_state of every output (io.cpp:91-159): *nout receives the file's
 * output count and key_out (1024 bytes) the key of output number index in
 * file order (DGX_ERR_INVALID when index >= *nout). */
dgx_status dgx_read_dg_output_key(const char* path, uint64_t index, char* key_out,
                                  uint64_t* nout);

/* The simstate of the output named key (DGX_ERR_IO when absent); out holds
 * nelem*5*(p+1)^3 doubles sized from dgx_read_dg_header. */
dgx_status dgx_read_dg_output(const char* path, const char* key, double* out);

#ifdef __cplusplus
}
#endif

#endif /* DGX_H */
