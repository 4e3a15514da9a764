/*
 * swm_gpu.h -- C-ABI of the MI355X (gfx950) SWM time-stepping path.
 *
 * The reference has no FFI layer: its "interface" is the set of Fortran module
 * procedures that src/model.f90 calls in a fixed order, plus module arrays.
 * Each entry point below replaces one group of them; a Fortran
 * ISO_C_BINDING shim keeping the reference procedure names calls these
 * (swm_amd/fortran/swm_gpu_shim.f90).
 *
 * Conventions
 *  - Every array argument is a host pointer to a Fortran (Nx,Ny) array,
 *    leading dimension Nx (C view [Ny][Nx]); (Nx,Ny,3) AB history arrays are
 *    three such planes back to back.  Masks are int8 0/1 "ocean" masks.
 *  - Host memory stays caller-owned and is only read/written inside the call.
 *    Device memory is owned by the opaque context.
 *  - Return value 0 = success; otherwise a non-zero status and
 *    swm_gpu_last_error() gives the message.  No C++ exception crosses the ABI.
 *  - One host thread per context; all device work is ordered on the context's
 *    own stream; swm_gpu_sync/_sor/_get block until their results are on the host.
 *  - The HIP path is the only path: there is no CPU fallback.  A missing or
 *    unusable GPU makes swm_gpu_create fail.
 */
#ifndef SWM_GPU_H
#define SWM_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_GPU_ABI_VERSION 2

/* model.h physics variant (include/model.h:9-12 of each experiment) */
enum swm_variant {
    SWM_FULLY_NONLINEAR = 1,
    SWM_LINEARISED_MEAN_STATE = 2,          /* needs u_bs, v_bs, zeta_bs and Ny <= Nx */
    SWM_LINEARISED_STATE_OF_REST = 3
};
/* model.h switches */
enum swm_flags {
    SWM_LATERAL_MIXING = 1,                 /* LATERAL_MIXING */
    SWM_QUADRATIC_BOTTOM_FRICTION = 2,      /* QUADRATIC_BOTTOM_FRICTION */
    SWM_BAROTROPIC = 4,                     /* BAROTROPIC */
    SWM_TAU_SCALE = 8,                      /* TAU_SCALE defined (value in swm_desc.tau_scale) */
    SWM_ELLIPTIC_SOLVER = 16,               /* built with CALC_LIB_ELLIPTIC_SOLVER = ElSolv_SOR (include/calc_lib.h:7-18) */
    SWM_FP32_LOW_PART = 32                  /* no model.h counterpart: an SWM_FP32 context carries a bfloat16
                                               low part of u, v, eta and CH (32 significant bits of state,
                                               4 B more per field and step); ignored in FP64 */
};
enum swm_precision { SWM_FP64 = 0, SWM_FP32 = 1 };
enum swm_grid_id { SWM_GRID_H = 0, SWM_GRID_U = 1, SWM_GRID_V = 2, SWM_GRID_ETA = 3 };
enum swm_component { SWM_COMP_X = 0, SWM_COMP_Y = 1, SWM_COMP_ETA = 2 };
/* swm_forcing_nl forcingtype (src/swm_forcing_module.f90:223-243) of a
 * time-interpolated stream */
enum swm_stream_kind { SWM_STREAM_WIND = 0, SWM_STREAM_CUSTOM = 1, SWM_STREAM_HEATING = 2 };

/* grid_t, src/grid_module.f90:9-20 (H, ocean, cos_lat, tan_lat, f) */
typedef struct swm_grid_desc {
    const int8_t *ocean;     /* (Nx,Ny) ocean mask; land = 1 - ocean */
    const double *H;         /* (Nx,Ny) layer thickness on this grid as left by init */
    const double *cos_lat;   /* (Ny) */
    const double *tan_lat;   /* (Ny) */
    const double *f;         /* (Ny) Coriolis parameter */
} swm_grid_desc;

/* Everything the step reads that the reference computes at init:
 * initDomain (src/domain_module.f90:49-207), SWM_damping_init
 * (src/swm_damping_module.f90:44-117), SWM_forcing_init
 * (src/swm_forcing_module.f90:78-127), time_integration_init
 * (src/time_integration_module.f90:35-52), vars_module constants
 * (src/vars_module.f90:27-40).  NULL array pointers take the reference's
 * initVar defaults (impl = 1, gamma = 0, forcing = 0). */
typedef struct swm_desc {
    int32_t abi_version;     /* SWM_GPU_ABI_VERSION */
    int32_t Nx, Ny;
    int32_t variant;         /* enum swm_variant */
    int32_t flags;           /* enum swm_flags */
    int32_t precision;       /* enum swm_precision */
    double G, Ah, dt, A, dLambda, dTheta, RHO0, minD, lbc, ab_chi, tau_scale;
    swm_grid_desc grid[4];   /* indexed by enum swm_grid_id */
    const double *impl_u, *impl_v, *impl_eta;           /* implicit damping divisors */
    const double *gamma_sq_u, *gamma_sq_v;              /* quadratic friction */
    const double *F_x_const, *F_y_const, *F_eta_const;  /* constant forcing */
    /* windstress streams re-applied every step as tau/(RHO0*Du) (FULLY_NONLINEAR,
     * src/swm_forcing_module.f90:274-324,631-632); NULL = none */
    const double *tau_x, *tau_y;
    /* oscillating forcing per component (src/swm_forcing_module.f90:326-374);
     * NULL = none.  Applied after windstress. */
    const double *osc_re[3], *osc_im[3];
    double osc_omega[3];
    /* LINEARISED_MEAN_STATE basic state as SWM_timestep_init leaves it
     * (src/swm_timestep_module.f90:76-87): u_bs, v_bs (evaluateStreamfunction of
     * psi_bs) and zeta_bs (laplacian of psi_bs on the H grid); NULL otherwise */
    const double *u_bs, *v_bs, *zeta_bs;
} swm_desc;

/* TRC_tracer, src/tracer_vars.f90:36-50 */
typedef struct swm_tracer_desc {
    double kappa_h;          /* horizontal diffusivity */
    const double *CH;        /* (Nx,Ny) layer-integrated tracer at level N0 */
    const double *C;         /* (Nx,Ny) concentration as held by the host (read by the next step) */
    const double *G_CH;      /* (Nx,Ny,3) AB history, NULL = zero */
    const double *C0;        /* restoring field, NULL = 0 */
    const double *gamma_C;   /* relaxation coefficient, NULL = 0 */
    const double *impl;      /* 1 + dt*cons, NULL = 1 */
} swm_tracer_desc;

typedef struct swm_ctx swm_ctx;

/* Replaces SWM_initSWM / initCalcLib / init_ElSolv_SOR set-up of device data
 * (src/swm_module.f90:37-48, src/calc_lib.f90:133-176, src/ElSolv_SOR.f90:44-66):
 * uploads static fields, builds coefficient tables, allocates device state. */
int swm_gpu_create(const swm_desc *d, swm_ctx **out);

/* State hand-over after SWM_initialConditions (src/swm_module.f90:156-181):
 * SWM_u/v/eta at N0 and G_u/G_v/G_eta (Nx,Ny,3) as held after SWM_advance.
 * G pointers may be NULL (zero history).  itt is vars_module::itt. */
int swm_gpu_upload_state(swm_ctx *c, const double *u, const double *v, const double *eta,
                         const double *G_u, const double *G_v, const double *G_eta, int64_t itt);

/* Replaces TRC_add_to_list (src/tracer_vars.f90:124-185) for one tracer.
 * Tracers are stepped in the order they were added (TRC_timestep list order). */
int swm_gpu_add_tracer(swm_ctx *c, const swm_tracer_desc *t, int32_t *tracer_id);

/* Replaces SWM_forcing_initStream (src/swm_forcing_module.f90:593-640) for a
 * dataset read through a memoryChunk (src/memchunk_module.f90:77-115): nrec
 * (Nx,Ny) records back to back, spaced rec_dt model seconds (getDt,
 * memchunk_module.f90:252-273), record 1 at model time 0 and the dataset
 * repeating after its last record; chunk_size is the namelist chunksize
 * (SWM_DEF_FORCING_CHUNKSIZE = 100).  Every step adds, on the component's
 * ocean points, the dataset interpolated in time as getChunkData(itt*dt)
 * does (214-243): kind WIND as [TAU_SCALE*]tau/(RHO0*D_face) (divided by RHO0
 * alone unless BAROTROPIC or FULLY_NONLINEAR; components X, Y), CUSTOM as is
 * (X, Y, ETA), HEATING as is (ETA).  Streams are processed after the
 * descriptor's windstress and oscillating streams, in the order added (the
 * reference's list order when the namelist lists them last).  All records stay
 * resident on the device (nrec planes).  At most 4 streams; a context with
 * streams is stepped alone (not in a batch) and without graph replay. */
int swm_gpu_add_forcing_stream(swm_ctx *c, int32_t kind, int32_t component, int32_t nrec, double rec_dt,
                               int32_t chunk_size, const double *records, int32_t *stream_id);

/* Replaces DFF_initDynFromFile (src/dynFromFile_module.f90:56-95) for one
 * DYNFROMFILE dataset: nrec (Nx,Ny) records of u (SWM_COMP_X), v (SWM_COMP_Y)
 * or eta (SWM_COMP_ETA) spaced rec_dt model seconds, read through a
 * memoryChunk of chunk_size records (DFF_Nt_chunksize).  At every
 * model_advance the master fields become ocean * getChunkData(itt*dt) +
 * SWM field (DFF_advance, 137-160; model.f90:238-259): the tracers are
 * advected by that flow, and swm_gpu_sync("U"/"V"/"ETA") returns the master
 * fields.  Streamfunction datasets (FileName_psi) are not taken.  A context
 * with datasets is stepped alone, on the tile kernels, without graph replay. */
int swm_gpu_add_dyn_stream(swm_ctx *c, int32_t field, int32_t nrec, double rec_dt, int32_t chunk_size,
                           const double *records);

/* nsteps iterations of the time loop body of src/model.f90:159-196 without Diag:
 * SWM_timestep (src/swm_module.f90:71-74), TRC_timestep
 * (src/tracer_module.f90:109-123) and model_advance (src/model.f90:238-259). */
int swm_gpu_step(swm_ctx *c, int64_t nsteps);

/* vars_module::itt as seen by the device state. */
int64_t swm_gpu_itt(const swm_ctx *c);

/* Copy one registered field to the host before a due diag task
 * (src/diagTask.f90:266-303 reads them through the register,
 * src/vars_module.f90:85-104).  Names are the reference's register names:
 * SWM_U SWM_V SWM_ETA (also U V ETA), G_U G_V G_ETA (newest slot),
 * SWM_D SWM_DH SWM_DU SWM_DV SWM_EDENS SWM_POT SWM_RELVORT SWM_MU SWM_MV
 * F_X F_Y F_ETA SWM_LATMIX_U SWM_LATMIX_V, and for tracer_id >= 0:
 * CH C G_CH uhc vhc.  Derived fields are materialised on demand from the
 * device state.  The diagnostic variables of src/diagVar.f90:115-134:
 *   PSI   computeStreamfunction (src/calc_lib.f90:727-780) of u, v at N0,
 *         on the device in FP64; with SWM_ELLIPTIC_SOLVER of the non-divergent
 *         flow u_nd, v_nd (as that build's computeStreamfunction does);
 *   U_ND, V_ND  computeNonDivergentFlowField (src/calc_lib.f90:548-590):
 *         u - d(chi)/dx, v - d(chi)/dy with SWM_ELLIPTIC_SOLVER, else u, v;
 *   CHI   with SWM_ELLIPTIC_SOLVER computeVelocityPotential
 *         (src/calc_lib.f90:500-525): main_ElSolv_SOR of the divergence of u,
 *         v with epsilon = EPS and first_guess = (itt == 1), chi kept on the
 *         device between steps, solved at most once per step; without it, the
 *         last swm_gpu_velocity_potential solution.
 * Blocks. */
int swm_gpu_sync(swm_ctx *c, const char *name, int32_t tracer_id, double *host_out);

/* main_ElSolv_SOR(B, chi, epsilon, first_guess) contract
 * (src/ElSolv_SOR.f90:150-222): red-black SOR with the reference's fixed
 * Chebyshev omega schedule and RMS-of-correction stopping rule.  B and chi are
 * host (Nx,Ny) arrays; chi is the first guess on entry and the result on exit.
 * iters = sweeps used; anorm = final sqrt(sum res^2).  Returns 0 when
 * converged, 1 when max_count sweeps did not converge (the reference only
 * warns), other values on error. */
int swm_gpu_sor(swm_ctx *c, const double *B, double *chi, double epsilon, int32_t first_guess,
                int64_t *iters, double *anorm);

/* Fixed-omega red-black SOR: stops when sqrt(sum r^2)/sqrt(sum B^2) < rel_tol
 * after a full sweep, r = (L chi - B) at each cell when it is updated, or after
 * max_sweeps.  fp32_sweeps != 0 runs the sweeps on FP32 copies of chi/B with
 * the residual norm and convergence test in FP64.  Same return convention. */
int swm_gpu_sor_fixed(swm_ctx *c, const double *B, double *chi, double omega, double rel_tol,
                      int64_t max_sweeps, int32_t fp32_sweeps, int64_t *iters, double *rel_res);

/* Averaging diag tasks kept on the device (processTask / addDataToTaskBuffer,
 * src/diagTask.f90:266-340): a registered buffer receives buffer += dt *
 * field**p (p = 2 for square averaging) in FP64 after every step, including
 * steps replayed from graphs, so averaging tasks need no per-step copies.
 * Fields: SWM_U, SWM_V, SWM_ETA, CH (tracer_id).  avg_read returns
 * buffer / bufferCount (bufferCount = the accumulated dt) and resets both when
 * reset != 0, as the reference does after writing an average.  Contexts with
 * averages are stepped alone (not in a batch). */
int swm_gpu_avg_add(swm_ctx *c, const char *name, int32_t tracer_id, int32_t square, int32_t *avg_id);
int swm_gpu_avg_read(swm_ctx *c, int32_t avg_id, double *host_out, double *weight, int32_t reset);

/* Tuning knobs and the autotuner.  No counterpart in the reference (its loops
 * have no launch shapes); they serve callers that pick kernel shapes for their
 * grids.  Shapes 0..swm_gpu_kernel_shape_count()-1: 0 = 64x4 LDS tiles,
 * 1 = 32x8 (default for FP32 grids above 2^22 cells), 2 = 64x8, 3 = 128x4,
 * 4 = 32x16, 5..8 = marching wavefronts (automatic, 8, 16, 4 rows), 9, 10 =
 * persistent 64x4, 32x8 tiles, 11, 12 = 32x8, 64x4 tiles loading the AB
 * history after the halo (11 is the default for every other grid); every
 * shape gives the same FP64
 * results bit for bit.  swm_gpu_tune times
 * `steps` steps of every shape from a device snapshot of the state, restores
 * the state exactly, selects the fastest and returns it (times_ms, if not
 * NULL, receives kernel_shape_count per-step times); with steps >= 6 it then
 * times the best shape's steps replayed from captured six-step graphs against
 * direct launches and keeps the faster launch form (swm_gpu_launch_mode reads
 * it and both per-step times; -1 for a form not timed).  set_launch_mode
 * picks it by hand (1: graphs, 0: direct launches, the default).
 * set_sor_check_every makes the fixed-omega SOR test convergence every `every`
 * sweeps (1..64; the reference mode always tests every sweep).
 * set_sor_halo sets the LDS halo of the blocked fixed-omega sweeps in cells:
 * 2K cells for K sweeps per launch (2, 4 or 8; K must divide the cadence),
 * 0 for one launch per colour, -1 (default) for the largest of 4, 2, 1 sweeps
 * that divides the cadence.  The reference mode (test after every sweep) uses
 * K = halo / 2 sweeps per launch, 4 for halo -1 (each launch also keeps the
 * chi and residual of its earlier sweeps, so any of its sweeps can be the one
 * that passes), and one launch per colour for halo 0.  Every halo gives the
 * same chi bit for bit and, in the reference mode, the same sweep count. */
int swm_gpu_kernel_shape_count(void);
int swm_gpu_set_kernel_shape(swm_ctx *c, int32_t shape);
int swm_gpu_tune(swm_ctx *c, int32_t steps, int32_t *best_shape, double *times_ms);
int swm_gpu_set_launch_mode(swm_ctx *c, int32_t graphs);
int swm_gpu_launch_mode(const swm_ctx *c, int32_t *graphs, double *ms_direct, double *ms_graph);
int swm_gpu_set_sor_check_every(swm_ctx *c, int32_t every);
int swm_gpu_set_sor_halo(swm_ctx *c, int32_t halo);
/* rows of the blocked sweeps' 64-column LDS tile: 16 (default) or 32 (fewer
 * halo cells per output; not with the 8-cell halo, whose 64x32 tile would
 * exceed 64 KiB of LDS).  Every tile gives the same chi bit for bit. */
int swm_gpu_set_sor_tile_rows(swm_ctx *c, int32_t rows);

/* computeVelocityPotential (src/calc_lib.f90:500-525) without host copies:
 * B = computeDivergence of the current device velocity (calc_lib.f90:607-622)
 * in the sweep precision, then fixed-omega red-black SOR as in
 * swm_gpu_sor_fixed (FP64 ||B|| and convergence test also for FP32 sweeps),
 * warm-started from the previous solution of the same sweep precision kept
 * on the device (zero on the first call or when reset != 0).  The solution
 * stays on the device; swm_gpu_sync(c, "CHI", -1, out) copies the last one.
 * Same return convention as swm_gpu_sor_fixed. */
int swm_gpu_velocity_potential(swm_ctx *c, double omega, double rel_tol, int64_t max_sweeps, int32_t fp32_sweeps,
                               int32_t reset, int64_t *iters, double *rel_res);

/* Sweeps used and final sqrt(sum res^2) of the last computeVelocityPotential
 * solve of the elliptic-solver build (the reference logs them at debug level,
 * src/ElSolv_SOR.f90:213-220). */
int swm_gpu_elliptic_stats(swm_ctx *c, int64_t *iters, double *anorm);

/* computeDivergence(u, v, div, u_grid, v_grid) of the current device velocity
 * (src/calc_lib.f90:607-622), the right-hand side computeVelocityPotential
 * hands to the SOR (src/calc_lib.f90:500-525).  Copies div to the host. */
int swm_gpu_divergence(swm_ctx *c, double *host_div);

/* Conservation sums of the current device state in FP64 with compensated
 * (Neumaier) accumulation, also for FP32 storage.  No counterpart in the
 * reference, which has no conservation diagnostics; it serves callers that
 * check an FP32 run's mass and energy drift against an FP64 one.  Over ocean
 * eta points, weighted by cos(lat_eta):
 *   out[0] = sum eta
 *   out[1] = sum 0.5*D*(0.5*(u(i)^2+u(i+1)^2) + 0.5*(v(j)^2+v(j+1)^2)) + 0.5*G*eta^2
 *   out[2] = sum CH of tracer tracer_id (0 when tracer_id < 0)
 * D = eta + H_eta for FULLY_NONLINEAR, H_eta otherwise.  Multiply by
 * A^2*dLambda*dTheta for volumes.  Blocks until the result is on the host. */
int swm_gpu_sums(swm_ctx *c, int32_t tracer_id, double *out);

/* The same sums as swm_gpu_sums without any device allocation: no
 * counterpart in the reference; it serves callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Its partial sums, result and cos(lat) rows live in
 * a static device array of the library (one per GPU, about 180 KB, held
 * until the process ends), so grids of at most 16384 rows (Ny) are accepted;
 * larger ones fail with status 4.  Calls that overlap on one GPU (from
 * several host threads or contexts at once) share that array and race. */
int swm_gpu_sums_noalloc(swm_ctx *c, int32_t tracer_id, double *out);

/* Batched ensembles.  No counterpart in the reference, which advances one
 * domain per process (model.f90:159-196 per member); these serve callers that
 * run many independent domains of uneven size at once (ensembles, a request
 * server).  Members are ordinary contexts, all of one precision and variant,
 * without oscillating forcing.  Each batch step is one launch covering every
 * tile of every attached member (SWM_timestep + TRC_timestep + model_advance
 * per member, the same arithmetic as swm_gpu_step, so a member's result is the
 * same alone and in the batch) plus a launch that moves the members' step
 * state on; the pair is captured once in a hipGraph and replayed.
 *   attach:    the context joins and will advance nsteps more steps; while
 *              attached, swm_gpu_step / upload_state / add_tracer on it fail,
 *              and sync / sums / divergence / itt see its current state.
 *   step:      advances every member with steps left by up to nsteps steps
 *              (members stop at their own count); *active (optional) = members
 *              with steps left afterwards.
 *   detach:    the context leaves with its state; set_steps resets a member's
 *              remaining count; remaining reads it. */
typedef struct swm_batch swm_batch;
int swm_gpu_batch_create(swm_batch **out);
int swm_gpu_batch_attach(swm_batch *b, swm_ctx *c, int64_t nsteps);
int swm_gpu_batch_detach(swm_batch *b, swm_ctx *c);
int swm_gpu_batch_set_steps(swm_batch *b, swm_ctx *c, int64_t nsteps);
int swm_gpu_batch_remaining(swm_batch *b, const swm_ctx *c, int64_t *remaining);
int swm_gpu_batch_step(swm_batch *b, int64_t nsteps, int32_t *active);
/* batch tile and launch form, 0..swm_gpu_batch_shape_count()-1: 0 = 64x4 and
 * 1 = 32x8 (default) tiles with one block per tile; 2 = 64x4 and 3 = 32x8
 * tiles drained by a persistent grid of resident blocks from per-XCD tile
 * queues (one launch per step either way; every form gives the same results
 * bit for bit).  batch_tune times `steps` steps of each from a device snapshot
 * of every member, restores the members exactly, selects the fastest
 * (times_ms: shape_count per-step times, may be NULL). */
int32_t swm_gpu_batch_shape_count(void);
int swm_gpu_batch_set_shape(swm_batch *b, int32_t shape);
int swm_gpu_batch_tune(swm_batch *b, int32_t steps, int32_t *best_shape, double *times_ms);
/* device time of the last swm_gpu_batch_step (ms) */
int swm_gpu_batch_last_timing(const swm_batch *b, double *total_ms);
const char *swm_gpu_batch_last_error(const swm_batch *b);
void swm_gpu_batch_destroy(swm_batch *b);

/* Time spent in the last swm_gpu_step call's kernels (ms, device clock) and
 * the average duration of the step kernel (ms), for roofline accounting. */
int swm_gpu_last_timing(const swm_ctx *c, double *total_ms, double *step_kernel_ms);

const char *swm_gpu_last_error(const swm_ctx *c);
void swm_gpu_destroy(swm_ctx *c);

#ifdef __cplusplus
}
#endif
#endif /* SWM_GPU_H */
