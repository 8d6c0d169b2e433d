/*
 * vi_mi355x.h -- C ABI of the MI355X (gfx950) viscoelastic velocity-stress stepper.
 *
 * The library replaces the time-stepping path of haralddd/ViscoelasticInversion:
 *   Parameters(; Nx, Nz, dx, dz, dt, ..., model, cpml, free_surface, source)
 *       src/sciml_solver/Parameters.jl:47-86                -> vi_create + vi_set_*
 *   make_problem(p; s0, v0, tspan)  ViscoelasticProblem.jl:273-292
 *                                                           -> vi_set_state
 *   solve_problem(prob; ...)        ViscoelasticProblem.jl:307-309
 *   leapfrog_step!(fields, parameters, dt)  src/visco-kernels.jl:87-100
 *                                                           -> vi_step / vi_sync
 *   _vel_1Q! / _stress_1Q!          src/visco-kernels.jl:3-20, 59-85   (fused kernels)
 *   RickerSource / inject_source_kernel!  src/sciml_solver/Source.jl:7-58 -> vi_set_sources
 *   CPMLConfig / CPMLCoefficients   src/sciml_solver/CPML.jl:28-221     -> vi_set_cpml
 *   FreeSurfaceBC / apply_free_surface!   src/sciml_solver/CPML.jl:366-393 -> vi_set_free_surface
 *   H-AFDA surface: FreeSurfaceStencil, Stencil.jl:52-105, and the row kernels of
 *                   free-surface-kernels.jl:16-214      -> vi_set_free_surface(h, 2),
 *                                                          vi_set_afda_points, vi_afda_weights
 *   H-AFDA formula #3: _stencil_hermitian, Stencil.jl:119-189 (D3, D3_deriv of
 *                   Stencil.jl:52-105), _stress_freesurf_j2! free-surface-kernels.jl:154-214
 *                                                       -> vi_set_free_surface(h, 3),
 *                                                          vi_set_afda_hermitian,
 *                                                          vi_afda_hermitian_weights
 *   IsotropicModel / QModel         src/sciml_solver/Models.jl:3-25, 90-107 -> vi_set_model,
 *                                                                              vi_set_relaxation
 *   VTIModel                        src/sciml_solver/Models.jl:27-33   -> vi_set_model_vti
 *
 * Conventions
 *   - Every 2D host array is [nz][nx] with x fastest (Julia's column-major (Nx, Nz)).
 *   - Host arrays are in the handle's precision (float for VI_F32, double for VI_F64).
 *     The library copies them during the call and never keeps host pointers.
 *   - All functions return 0 on success or a negative VI_E* code; vi_last_error() then
 *     returns a message (thread-local). No C++ exception crosses this boundary.
 *   - A handle owns its device memory and one HIP stream. vi_step is asynchronous until
 *     vi_sync or any vi_get_* call. A handle is used by one host thread at a time.
 */
#ifndef VI_MI355X_H
#define VI_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vi_sim_s* vi_handle;
typedef struct vi_model_s* vi_model;
typedef struct vi_batch_s* vi_batch;
typedef struct vi_grad_s* vi_grad;

enum { VI_F32 = 0, VI_F64 = 1 };

enum {
  VI_OK = 0,
  VI_EINVAL = -1,   /* bad argument */
  VI_ESTATE = -2,   /* call out of order (e.g. step before set_model) */
  VI_EDEVICE = -3,  /* HIP runtime error, no device, or kernel image missing */
  VI_ENOMEM = -4    /* device allocation failed */
};

/* fields for vi_get_state / vi_set_state (values match the oracle's numbering) */
enum {
  VI_VX = 0, VI_VZ = 1, VI_SXX = 2, VI_SZZ = 3, VI_SXZ = 4,
  VI_MXX = 5, VI_MZZ = 6, VI_MXZ = 7,                  /* [L][nz][nx] */
  VI_PSI_VX_X = 8, VI_PSI_VZ_X = 9, VI_PSI_SXX_X = 10, VI_PSI_SXZ_X = 11, /* [nz][2T] */
  VI_PSI_VX_Z = 12, VI_PSI_VZ_Z = 13, VI_PSI_SXZ_Z = 14, VI_PSI_SZZ_Z = 15, /* [2T][nx] */
  VI_NFIELDS = 16
};

/* source kinds: force_z is the reference's RickerSource (added to dvz before buoyancy,
 * ViscoelasticProblem.jl:162-171); explosive adds to the sxx and szz rates. */
enum { VI_SRC_FORCE_Z = 0, VI_SRC_EXPLOSIVE = 1, VI_SRC_FORCE_X = 2 };

/* receiver components */
enum { VI_REC_VX = 0, VI_REC_VZ = 1, VI_REC_SXX = 2, VI_REC_SZZ = 3, VI_REC_SXZ = 4 };

/* Grid and physics of one simulation (Parameters.jl:47-63: Nx, Nz, dx, dz, dt,
 * fd_order_x, fd_order_z). Staggered operators of order 2, 4, 6 or 8 per axis
 * (Stencil(xorder, zorder, dx, dz; stagger), Stencil.jl:280-320); the Dirichlet ring is 4
 * cells for every order (the reference pads by order/2). */
typedef struct {
  int nx, nz;
  double dx, dz, dt;
  int precision;       /* VI_F32 or VI_F64 */
  int nrelax;          /* L standard linear solids, 0 = elastic, at most 6 */
  int fd_order_x, fd_order_z; /* 2, 4, 6 or 8 */
} vi_grid_desc;

/* CPMLConfig (CPML.jl:28-63). d0 must be given explicitly; vi_cpml_auto_d0 computes the
 * Collino & Tsogka value the reference uses when d0 === nothing (CPML.jl:55-61). */
typedef struct {
  int thickness;
  double d0, alpha_max, kappa_max, power;
  int left, right, top, bottom;
} vi_cpml_desc;

/* ---- host-only helpers (no device needed) ----------------------------------------- */
const char* vi_last_error(void);
const char* vi_version(void);
/* staggered 8th-order weights / h, as _stencil (Stencil.jl:9-14) gives them */
void vi_fd_coeffs(double h, double out[4]);
/* staggered first-derivative weights / h of order 2, 4, 6 or 8 (_stencil at +-1/2 ..
 * +-(order-1)/2, Stencil.jl:9-14 and 280-320): out[l] multiplies u[c+l] - u[c-l-1] (D+) or
 * u[c+l-1] - u[c-l] (D-); entries beyond order/2 are 0 */
int vi_fd_weights(double h, int order, double out[4]);
/* H-AFDA one-sided d/dz weights / h (the FreeSurfaceStencil coefficients, Stencil.jl:52-105,
 * formulas #1, #2 and #4 of Kristek et al. 2002, carried to every row an 8th-order centred
 * stencil would take above the surface). Table rows: A0..A3 = d sxz/dz at z = k h (from
 * sxz(0) = 0 and sxz at (m+1/2) h), B0..B2 = d szz/dz at (k+1/2) h (szz at m h), C1..C3 =
 * d vz/dz at k h (vz at (m+1/2) h), D0..D2 = d vx/dz at (k+1/2) h (vx at m h); column m
 * multiplies grid row 4 + m. counts[row] = points used (A: 1..8 besides the boundary point,
 * others 2..8); NULL = the default {3,4,6,8, 3,4,6, 2,4,6, 3,4,6}. Unused columns are 0. */
int vi_afda_weights(double h, const int* counts, double out[13][8]);
int vi_afda_default_points(int out[13]);
/* H-AFDA formula #3 (the reference's _stencil_hermitian(pts_half, 0, 1, 1), Stencil.jl:119-189,
 * as FreeSurfaceStencil's D3 / D3_deriv, Stencil.jl:52-105): d vz/dz at z = h from d vz/dz(0)
 * and vz at the half points (m+1/2) h, m < points (2..8):
 *   d vz/dz(h) = out[0] * d vz/dz(0) + sum_m out[1 + m] * vz((m+1/2) h);
 * out[0] = alpha (dimensionless), out[1 + m] = beta_m / h, 0 for m >= points. Each weight is
 * the correctly rounded exact rational (then divided by h). */
int vi_afda_hermitian_weights(double h, int points, double out[9]);
/* d0 = -(power+1) vmax ln(R) / (2 thickness dx)   (CPML.jl:60) */
double vi_cpml_auto_d0(int thickness, double vmax, double dx, double reflection_coef,
                       double power);
/* a, b, kappa of CPMLCoefficients along one axis (CPML.jl:109-189), computed in the given
 * precision as the reference computes them for its Float32 / Float64 arrays (the values the
 * kernels use), returned as double */
int vi_cpml_profile(int n, const vi_cpml_desc* c, double dt, int axis /*0=x,1=z*/,
                    int precision, double* a, double* b, double* kappa);
/* number of visible HIP devices (0 when none); never fails */
int vi_device_count(void);
/* free and total device memory of the current device; free includes the buffers the library
 * keeps for reuse (below), which it frees when an allocation needs the memory */
int vi_mem_info(size_t* free_bytes, size_t* total_bytes);
/* reuse of large device buffers (no counterpart in the reference). Buffers of 64 MB or more
 * that a handle releases (vi_destroy, vi_trim, a gradient workspace that is replaced) are
 * kept, once their device is idle, for the next allocation on that device that they fit (at
 * most 1/4 larger than asked), up to a third of the device memory; runs of jobs with fresh
 * handles then do not hand tens of GB back to the runtime and ask for them again. An
 * allocation that fails frees the kept buffers of its device and is tried again. */
int vi_cached_bytes(size_t* bytes);  /* bytes kept on the current device */
int vi_release_cached(void);         /* free them */

/* ---- device-resident models (no counterpart in the reference: IsotropicModel arrays are
 * re-created per Parameters). A model is uploaded once and shared by any number of shots,
 * each reading a column window of it (multi-shot acquisition on one large model). ------ */
int vi_model_create(int nx, int nz, int precision, const void* rho, const void* lam,
                    const void* mu, const void* tau /* NULL for elastic */, vi_model* out);
/* VTI model: stiffnesses C11, C13, C33, C55 (VTIModel, Models.jl:27-33; the stress update
 * is the TTI form of ViscoelasticProblem.jl:21-35 with C15 = C35 = 0, with the same SLS
 * relaxation tau applied to every modulus) */
int vi_model_create_vti(int nx, int nz, int precision, const void* rho, const void* c11,
                        const void* c13, const void* c33, const void* c55,
                        const void* tau /* NULL for elastic */, vi_model* out);
/* releases the caller's reference; shots that use the model keep it alive */
int vi_model_destroy(vi_model m);

/* ---- single-shot stepper ----------------------------------------------------------- */
int vi_create(const vi_grid_desc* g, vi_handle* out);
int vi_destroy(vi_handle h);
/* rho, lam, mu (and tau when nrelax > 0) as [nz][nx]; tau may be NULL when nrelax == 0 */
int vi_set_model(vi_handle h, const void* rho, const void* lam, const void* mu,
                 const void* tau);
/* VTI medium (VTIModel, Models.jl:27-33): rho, C11, C13, C33, C55 (and tau) as [nz][nx] */
int vi_set_model_vti(vi_handle h, const void* rho, const void* c11, const void* c13,
                     const void* c33, const void* c55, const void* tau);
/* use columns [x0, x0 + nx) of a shared model (same nz and precision) */
int vi_set_model_window(vi_handle h, vi_model m, int x0);
/* taun[L] relaxation times of the SLS mechanisms (QModel.τns, Models.jl:56-62) */
int vi_set_relaxation(vi_handle h, const double* taun);
/* NULL disables the CPML; resets the psi memory to zero */
int vi_set_cpml(vi_handle h, const vi_cpml_desc* c);
/* traction-free surface on row 4 (sxx, szz = 0, vx on the surface; vz, sxz half a cell
 * below). mode 0: none (Dirichlet ring); 1: stress imaging, szz and sxz mirrored with opposite
 * sign into rows 0..3 (CPML.jl:372-380); 2: H-AFDA, the vertical derivatives of rows 4..7 come
 * from one-sided stencils on rows 4..11 (vi_afda_weights) and nothing above the surface is
 * read or written; 3: mode 2 with d vz/dz on row 5 (z = h) from the Hermitian formula #3
 * (vi_afda_hermitian_weights) whose d vz/dz(0) is the surface row's value that keeps
 * szz = 0 (-lam/(lam + 2 mu) d vx/dx when elastic, free-surface-kernels.jl:160-172; with the
 * memory terms when L > 0; -C13/C33 d vx/dx for VTI). Needs nz >= 17 and no top CPML strip.
 * The gradient entry points take every mode: the adjoint transposes the one-sided rows, and
 * with mode 3 row 5's dependence on the surface row's value. */
int vi_set_free_surface(vi_handle h, int mode);
/* points per H-AFDA stencil (see vi_afda_weights; NULL = default). All-8 counts, the full
 * order of the reference's FreeSurfaceStencil(8), diverge in heterogeneous media. */
int vi_set_afda_points(vi_handle h, const int* counts);
/* half points of formula #3 (2..8, default 5). The reference's FreeSurfaceStencil(8) uses 8;
 * 6, 7 and 8 grow without bound in heterogeneous media (measured), 2 equals mode 2's row. */
int vi_set_afda_hermitian(vi_handle h, int points);
/* n sources at interior cells (ix, iz); wavelets [n][nt] in the handle's precision,
 * already scaled by the source coefficient; sample k is used by step k */
int vi_set_sources(vi_handle h, int n, const int* ix, const int* iz, const int* kind, int nt,
                   const void* wavelets);
/* n receivers at interior cells; traces are recorded for up to nt_max steps after this call */
int vi_set_receivers(vi_handle h, int n, const int* ix, const int* iz, const int* comp,
                     int nt_max);
int vi_set_state(vi_handle h, int field, const void* host);
int vi_get_state(vi_handle h, int field, void* host);
/* zero every state array (v, s, M, psi), the trace buffer and the step counter */
int vi_reset(vi_handle h);
/* advance nsteps full steps (velocity kernel, then stress kernel, per step) */
int vi_step(vi_handle h, int nsteps);
int vi_sync(vi_handle h);
long vi_step_index(vi_handle h);
/* Right-hand side at the handle's current state, without changing it: the analogue of
 * stress_eq!(ds, s, v, p, t) and velocity_eq!(dv, s, v, p, t)
 * (src/sciml_solver/ViscoelasticProblem.jl:62-174) for the fused scheme. dv [2][nz][nx] =
 * (vx, vz) rates of the velocity half step from the stored stresses (force sources added
 * before the buoyancy); ds [3][nz][nx] = (sxx, szz, sxz) rates of the stress half step from
 * the stored velocities and memory variables (explosive sources added; dszz = 0 on the free-
 * surface row; with stress imaging the image rows carry the negated rates of their mirror
 * rows); dm [3][L][nz][nx] = memory-variable rates (NULL to skip). src_values [nsrc] = the
 * sample of each source as given to vi_set_sources (NULL = no source term). CPML strips use
 * psi' = b psi + a d as the step does (psi is not changed), so a half step of vi_step is
 * x + dt * rate in the handle's precision. 0 outside the interior. The device workspace
 * (rate planes) is allocated on the first call and kept by the handle. */
int vi_rates(vi_handle h, const void* src_values, void* dv, void* ds, void* dm);
/* vi_rates without any device allocation (no counterpart in the reference). For callers that
 * cannot allocate device memory: an external time integrator evaluating the right-hand side
 * in a tight loop, serving loops that keep the device footprint fixed. The workspace is a
 * static device array of 128 MiB (5 + 3 L rate planes of the padded grid plus the source
 * list; larger handles are refused with VI_EINVAL), one per GPU, held until the process
 * ends. Calls that overlap on one GPU share it and race; the results of a call do not depend
 * on earlier calls. */
int vi_rates_noalloc(vi_handle h, const void* src_values, void* dv, void* ds, void* dm);
/* *finite = 1 when every interior velocity value is finite, 0 after a blow-up (an unstable
 * time step or model); waits for the handle's queued work. No counterpart in the reference:
 * a serving loop fails a diverging job alone with it. */
int vi_check_finite(vi_handle h, int* finite);
/* traces recorded so far: out[nrec][nt] for steps [0, nt) since vi_set_receivers */
int vi_get_traces(vi_handle h, void* out, int nt);
/* device bytes owned by the handle (state, CPML memory, traces, tables, private model and
 * the workspaces it keeps: see vi_trim) */
int vi_device_bytes(vi_handle h, size_t* bytes);
/* free the workspaces a handle keeps for its next call (no counterpart in the reference): a
 * gradient leaves its checkpoint, slot and adjoint buffers with the handle (vi_gradient_bytes
 * of its largest call) so that repeated gradients reuse them, vi_rates its rate planes */
int vi_trim(vi_handle h);

/* ---- gradient by the discrete adjoint (no counterpart in the reference: Enzyme is a
 * dependency but never called; the ENGD optimiser's grad_fn, EnergyNaturalGradient.jl:246,
 * is the intended caller). J = 1/2 sum_t sum_r (d_r(t) - dobs_r(t))^2 over nt steps from a
 * zero initial state with the handle's sources and receivers (capacity >= nt); gradients of J
 * with respect to the handle's rho, lam, mu, tau arrays ([nz][nx], handle precision; any may
 * be NULL). Checkpoints every checkpoint_every steps (0 = ceil(sqrt(nt))) and recomputes each
 * segment but the last (recorded during the forward pass), so device memory is about
 * 2 sqrt(nt) states. checkpoint_every >= nt stores every step's forward state during the
 * forward pass and recomputes nothing (memory for nt steps; vi_gradient_bytes tells the
 * size). The results do not depend on checkpoint_every. The handle is left reset (zero
 * state, step 0) and keeps the workspace for its next gradient (vi_trim frees it). */
int vi_gradient(vi_handle h, int nt, const void* dobs /*[nrec][nt]*/, int checkpoint_every,
                double* misfit, void* g_rho, void* g_lam, void* g_mu, void* g_tau);
/* the same for a VTI model (VTIModel, Models.jl:27-33): gradients with respect to rho, C11,
 * C13, C33, C55 and tau */
int vi_gradient_vti(vi_handle h, int nt, const void* dobs /*[nrec][nt]*/, int checkpoint_every,
                    double* misfit, void* g_rho, void* g_c11, void* g_c13, void* g_c33,
                    void* g_c55, void* g_tau);
/* The same computation as a job that advances in chunks without blocking the caller, so one
 * host thread can interleave several gradients with other device work (serving).
 * begin allocates the job's buffers (vi_gradient_bytes) and copies dobs to the device before
 * it returns (the caller may free it then); the handle then belongs to the job (other calls
 * on it fail with VI_ESTATE) until end or cancel.
 * advance enqueues at most max_steps forward, recomputed or transposed steps on the handle's
 * stream and returns without waiting (the residual and the misfit are computed on the
 * device); *finished = 1 once all work is enqueued. idle returns 1 when everything enqueued so far has completed, 0 otherwise.
 * end finishes any remaining work, waits, writes the misfit and the gradients (isotropic:
 * rho, lam, mu, tau; VTI: rho, C13, C55, tau, C11, C33; any may be NULL), resets the handle
 * and frees the job; cancel frees it without results. */
int vi_gradient_begin(vi_handle h, int nt, const void* dobs /*[nrec][nt]*/, int checkpoint_every,
                      vi_grad* out);
int vi_gradient_advance(vi_grad g, int max_steps, int* finished);
int vi_gradient_idle(vi_grad g);
int vi_gradient_end(vi_grad g, double* misfit, void* g_rho, void* g_lam_or_c13,
                    void* g_mu_or_c55, void* g_tau, void* g_c11, void* g_c33);
int vi_gradient_cancel(vi_grad g);

/* ---- device-memory planning (no counterpart in the reference; used for admission by an
 * HBM budget when many shots share one GPU) ---------------------------------------------- */
/* temporary device bytes one vi_gradient(h, nt, ..., checkpoint_every, ...) call allocates */
int vi_gradient_bytes(vi_handle h, int nt, int checkpoint_every, size_t* bytes);
/* upper bound of the device bytes of a shot before it is created: grid, model (VTI or not),
 * CPML strips, sources, receiver traces and, when grad_nt > 0, one gradient call; needs no
 * device */
int vi_plan_bytes(const vi_grid_desc* g, int vti, int cpml_thickness, int nsrc, int src_nt,
                  int nrec, int rec_nt, int grad_nt, int checkpoint_every, size_t* bytes);

/* ---- tiling (no counterpart in the reference: KA picks workgroup shapes itself) ------ */
/* tiling: tiles of 64 x (4 rows_per_wave) cells, rows_per_wave 1, 2, 4 or 8 (results are
 * identical); the 8- and 16-row tiles (2, 4) run as 8 wavefronts of one and two rows. Default:
 * 8-row tiles up to 2048 4-row tiles, 16-row tiles from 12288, 4-row tiles in between; a
 * batch defaults to 4-row tiles */
int vi_set_tiling(vi_handle h, int rows_per_wave);
int vi_get_tiling(vi_handle h);
/* time every variant for trial_steps steps on a snapshot of the state, restore the state,
 * keep the fastest; us_per_step (optional, 4 entries) receives the timings of 1, 2, 4, 8 */
int vi_autotune(vi_handle h, int trial_steps, int* chosen, double* us_per_step);

/* ---- batches: several shots advanced by one launch per half step -------------------
 * (no counterpart in the reference, which has one source per Parameters, Parameters.jl:59)
 * Members keep their own state; a batch launch covers all their tiles (grouped by precision
 * and nrelax). Every member's result is identical to stepping it alone. */
int vi_batch_create(vi_batch* out);
int vi_batch_destroy(vi_batch b);
/* set the members (replaces the previous set); handles must stay alive while members */
int vi_batch_set(vi_batch b, const vi_handle* hs, int n);
int vi_batch_set_tiling(vi_batch b, int rows_per_wave);
int vi_batch_get_tiling(vi_batch b);
/* cache blocking over shots (no counterpart in the reference): the shots of a precision/L
 * class are packed (first fit by decreasing working set) into groups whose working set
 * (state, memory variables, CPML memory, model columns) is at most `bytes` per stream share
 * (bytes / n streams), and each group runs all steps of a block (a
 * graph block, or the steps of one vi_batch_step call) before the next starts, so its data
 * stays in the Infinity Cache. Results are unchanged. 0 = every shot in every launch; the
 * default is 224 MB (of the 256 MB Infinity Cache). */
int vi_batch_set_cache_block(vi_batch b, size_t bytes);
size_t vi_batch_get_cache_block(vi_batch b);
/* n = 1..4 streams (default 2): the launch groups take the streams in turn, so n
 * groups run at once (each at most 1/n of the cache block), overlapping each other's launch
 * ramp and drain; 1: one group at a time. Results are unchanged. */
int vi_batch_set_streams(vi_batch b, int n);
/* times trial_steps of every tiling (1, 2, 4, 8 rows per wavefront; us_per_step[4]) and, when a
 * cache block is set, lockstep at the fastest tiling; keeps the fastest and restores the state */
int vi_batch_autotune(vi_batch b, int trial_steps, int* chosen, double* us_per_step);
/* replay a captured hipGraph of steps_per_graph steps (0 = direct launches) */
int vi_batch_set_graph(vi_batch b, int steps_per_graph);
/* advance every member by nsteps; asynchronous until vi_batch_sync or a member's vi_get_* */
int vi_batch_step(vi_batch b, int nsteps);
int vi_batch_sync(vi_batch b);
int vi_batch_profile_enable(vi_batch b, int every);
int vi_batch_profile_read(vi_batch b, int kernel, double* ms_total, long* launches);

/* ---- kernel timing: HIP events recorded by the launches themselves
 * (hipExtLaunchKernelGGL) on the handle's stream, for the launches of every `every`-th
 * step (1 = all; steps counted across calls since enabling); 0 disables. Enabling resets
 * the sums. -------------------------------------------------------------------------- */
int vi_profile_enable(vi_handle h, int every);
/* kernel 0 = velocity, 1 = stress: summed milliseconds and launch count since enable */
int vi_profile_read(vi_handle h, int kernel, double* ms_total, long* launches);
/* algorithmic HBM bytes per launch of kernel 0/1 for this handle's configuration */
int vi_algorithmic_bytes(vi_handle h, int kernel, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
