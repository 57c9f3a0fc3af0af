/*
 * pmgpu.h -- C ABI of the MI355X (gfx950) PM / P3M force step.
 *
 * This is the drop-in boundary: a host library that wants the GPU path of the
 * reference's PMMethod / P3MMethod / Grid / FFTAdapter / leapfrog (the C++
 * classes in include/pmsim/ are one such host, Python ctypes is another)
 * binds exactly these symbols.  Plain pointers and sizes only; no C++ types
 * and no exceptions cross it.
 *
 * Units: every quantity at this boundary is in the reference's code units
 * (include/unitConversions.h:8-46 of the reference): positions in grid
 * cells, velocities in cells per step, dt = 1, masses multiplied by
 * 4*pi*G*DT^2/H^3, so that lap(phi) = rho and a = -grad(phi).
 *
 * Particle arrays handed in and out are host arrays in the caller's particle
 * order ((x,y,z) interleaved, n*3 floats); the library keeps its own
 * device-resident, cell-sorted copy and maps back on output.
 *
 * Threading: a context owns one HIP stream.  A context is not re-entrant;
 * distinct contexts are independent.  Calls enqueue work on the context's
 * stream and return; the getters synchronise.
 *
 * Every call returns PMGPU_OK (0) or an error code; pmgpu_last_error() gives
 * the message.
 */
#ifndef PMGPU_H
#define PMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define PMGPU_OK 0
#define PMGPU_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument in the reference */
#define PMGPU_ERR_NOT_READY 2        /* e.g. a step before particles / G are set */
#define PMGPU_ERR_DEVICE 3           /* a HIP runtime error */
#define PMGPU_ERR_OUT_OF_MEMORY 4
#define PMGPU_ERR_UNSUPPORTED 5      /* valid in the reference, not on this path */

/* ---- enums: same order as the reference's enum classes --------------- */
/* InterpolationScheme, include/pmConfig.h:3 */
#define PMGPU_NGP 0
#define PMGPU_CIC 1
#define PMGPU_TSC 2
/* An extension: TSC about the nearest grid point (Hockney & Eastwood's
   assignment, whose weights are never negative).  PMGPU_TSC keeps the
   reference's anchor at the floor cell (source/pmMethod.cpp:250-269), whose
   centre weight 3/4 - d^2 is negative for d > sqrt(3)/2.  Same stencil,
   Green's functions and errors otherwise. */
#define PMGPU_TSC_CENTRED 3
/* FiniteDiffScheme, include/pmConfig.h:4 */
#define PMGPU_TWO_POINT 0
#define PMGPU_FOUR_POINT 1
/* GreensFunction, include/pmConfig.h:5 */
#define PMGPU_DISCRETE_LAPLACIAN 0
#define PMGPU_S1_OPTIMAL 1
#define PMGPU_S2_OPTIMAL 2
#define PMGPU_POOR_MAN 3
/* CloudShape, include/greensFunctions.h:7 */
#define PMGPU_S1 0
#define PMGPU_S2 1
/* external field kinds (the reference takes a std::function; the two
   kinds its demos use are a zero field and sphRadDecrField,
   source/externalFields.cpp:4-16) */
#define PMGPU_EXT_NONE 0
#define PMGPU_EXT_SPH_RAD_DECR 1

typedef struct pmgpu_ctx pmgpu_ctx;

/*
 * Construction parameters of the mesh part.  Replaces the PMMethod
 * constructor arguments (reference include/pmMethod.h:14-26) and the Grid
 * constructor (include/grid.h:12): grid points, interpolation, finite
 * difference, Green's function, particleDiameter (here already in code
 * units, as PMMethod stores it, source/pmMethod.cpp:56) and the effective
 * box (code units) used by escapedComputationalBox (pmMethod.cpp:146-156);
 * a box component <= 0 disables that test on its axis (periodic use: the
 * deposit and the interpolation always wrap indices periodically).
 * Grid dimensions must be powers of two in [32, 1024].
 */
typedef struct pmgpu_pm_params {
  int32_t grid[3];
  int32_t interpolation;
  int32_t finite_diff;
  int32_t greens_function;
  float particle_diameter;
  float box[3];
  int32_t ext_kind;      /* PMGPU_EXT_* */
  float ext_params[6];   /* sphRadDecrField: center x,y,z, R, M, G (original units) */
  float H, DT;           /* original-unit scales, for the external field only */
} pmgpu_pm_params;

/*
 * P3M short-range parameters.  Replaces the P3MMethod constructor
 * (reference include/p3mMethod.h:13-24, source/p3mMethod.cpp:19-48), all in
 * code units: cutoff radius re, particle diameter a, softening length,
 * cloud shape, useSRForceTable, and the computational box of the chaining
 * mesh.  enableYSorting only changes the reference's summation order and has
 * no counterpart here.
 * use_table: 0 = the direct formula (shortRangeForce, p3mMethod.cpp:220-238),
 *            1 = the reference's 500-entry table (useSRForceTable = true),
 *            2 = an accurate table of the pair factor in 128 log-spaced
 *                bins per binade of r^2 / rc^2 (cut from its float bits) over
 *                [2^-28, 1), within 2e-5 of the direct formula (an
 *                extension: the reference's table interpolates T ~ r^-3
 *                linearly in r^2 and is coarse for close pairs).
 */
typedef struct pmgpu_p3m_params {
  float cutoff;
  float particle_diameter;
  float softening;
  int32_t cloud_shape;
  int32_t use_table;
  float box[3];
} pmgpu_p3m_params;

/* ---- lifetime ------------------------------------------------------- */
/* Replaces Grid(gridPoints, fftAdapter) + PMMethod(...) construction. */
int pmgpu_create(const pmgpu_pm_params* params, pmgpu_ctx** out);
void pmgpu_destroy(pmgpu_ctx* ctx);
/* Message of the last failing call on ctx (ctx == NULL: of the last
   pmgpu_create on this thread). */
const char* pmgpu_last_error(const pmgpu_ctx* ctx);
/* Blocks until all work enqueued on the context's stream is done. */
int pmgpu_synchronize(pmgpu_ctx* ctx);

/* ---- particles ------------------------------------------------------- */
/* Uploads the particle state (replaces the PMMethod constructor's
   state/masses, source/pmMethod.cpp:57-59).  vel may be NULL (zero). */
int pmgpu_set_particles(pmgpu_ctx* ctx, int64_t n, const float* pos, const float* vel,
                        const float* mass);
/* Replaces PMMethod::getParticles (include/pmMethod.h:33): copies the
   state back in the caller's particle order.  Any pointer may be NULL. */
int pmgpu_get_particles(pmgpu_ctx* ctx, float* pos, float* vel, float* acc);
/* Adds accelerations evaluated by the host (code units, caller's particle
   order, n x 3) to those of the last force evaluation: the external field
   term of PMMethod::updateAccelerations (source/pmMethod.cpp:384-390) for a
   std::function the device cannot call.  Synchronises (acc is read before
   the call returns). */
int pmgpu_add_accelerations(pmgpu_ctx* ctx, const float* acc);
/* Overwrites positions / velocities / accelerations (caller order) without
   changing masses; NULL leaves a field as it is.  Lets a host that edited
   its AoS mirror push it back (PMMethod::getParticles returns a mutable
   reference in the reference). */
int pmgpu_update_particles(pmgpu_ctx* ctx, const float* pos, const float* vel,
                           const float* acc);
int64_t pmgpu_num_particles(const pmgpu_ctx* ctx);

/* ---- force step ------------------------------------------------------ */
/* PMMethod::initGreensFunction (source/pmMethod.cpp:164-185), evaluated in
   FP64 on the device; the stored influence function is the real,
   inversion-symmetrised G with the inverse-FFT 1/N folded in. */
int pmgpu_init_greens(pmgpu_ctx* ctx);
/* PMMethod::pmMethodStep (source/pmMethod.cpp:137-144): deposit -> R2C FFT
   -> x G(k) -> C2R FFT -> finite-difference field -> interpolation; writes
   the acceleration (plus the external field). */
int pmgpu_pm_step(pmgpu_ctx* ctx);
/* Configures the P3M short-range pass (P3MMethod constructor). */
int pmgpu_p3m_configure(pmgpu_ctx* ctx, const pmgpu_p3m_params* params);
/* P3MMethod::calculateShortRangeForces + correctAccelerations
   (source/p3mMethod.cpp:168-192, 50-57): adds the short-range correction
   to the acceleration of every particle. */
int pmgpu_p3m_short_range(pmgpu_ctx* ctx);
/* Record (on != 0) or stop recording the short-range part of each
   particle's acceleration in every short-range pass, and read the last
   pass's (code units, caller order, n x 3; zero for particles the pass did
   not reach).  The reference keeps it as shortRangeForce plus the
   shortRangeFromNeighbor slots of each Particle (include/particle.h:11-12,
   filled at source/p3mMethod.cpp:269-272, summed in correctAccelerations,
   :50-57); the recorded value is that sum divided by the mass.  Every pass
   writes every particle's value. */
int pmgpu_p3m_record_short_range(pmgpu_ctx* ctx, int32_t on);
int pmgpu_p3m_get_short_range(pmgpu_ctx* ctx, float* acc);

/* ---- leapfrog (source/leapfrog.cpp:5-24, dt in code units) ----------- */
#define PMGPU_HALF_KICK 0     /* setHalfStepVelocities: v += dt/2 a      */
#define PMGPU_INTEGER_STEP 1  /* setIntegerStepVelocities: vi = v + dt/2 a */
#define PMGPU_KICK 2          /* updateVelocities: v += dt a             */
#define PMGPU_DRIFT 3         /* updatePositions: x += dt v              */
int pmgpu_leapfrog(pmgpu_ctx* ctx, int32_t op, float dt);
/* setIntegerStepVelocities result (caller order, n*3), for diagnostics. */
int pmgpu_get_integer_step_velocities(pmgpu_ctx* ctx, float* vel);

/*
 * The arithmetic of PMMethod::run / P3MMethod::run (source/pmMethod.cpp:
 * 72-122, source/p3mMethod.cpp:73-154) without I/O or diagnostics: when
 * `initial` is nonzero, first the force and the half kick; then `iterations`
 * times: drift, escape test, force (+ short range when `p3m`), kick.
 * Kick and the next drift are fused on the device.  When a particle leaves
 * the box the remaining work is skipped on the device (the reference breaks
 * out of its loop); *iterations_done (may be NULL, then no sync) receives
 * the count of completed iterations.  With profiling off, iterations after
 * the first are replayed two at a time from a HIP graph captured on the
 * context's stream (kept until a call changes the step's arguments).
 */
int pmgpu_run(pmgpu_ctx* ctx, int32_t iterations, int32_t initial, int32_t p3m,
              int32_t* iterations_done);
/* escapedComputationalBox (pmMethod.cpp:151-156) of the current state. */
int pmgpu_escaped(pmgpu_ctx* ctx, int32_t* escaped);

/* GreenDiscreteLaplacian / GreenOptimal / GreenPoorMan
   (source/greensFunctions.cpp:122-220) at one unfolded mode (kx, ky, kz) in
   [0, N), evaluated on the host in FP64 with the device kernel's code (NaN
   for invalid arguments).  Unlike the FP32 reference it is exactly 0 where
   D(k) = 0 and uses a series where S1/S2Fourier cancel. */
double pmgpu_greens_value(int32_t greens_function, int32_t interpolation, int32_t finite_diff,
                          double particle_diameter, int32_t kx, int32_t ky, int32_t kz, int32_t nx,
                          int32_t ny, int32_t nz);

/* ---- diagnostics ------------------------------------------------------ */
/* FP64 device reductions of SimInfo (source/simInfo.cpp:50-118) on the
   current state, in code units, written to out[PMGPU_DIAG_COUNT]:
     [0] kinetic energy sum 1/2 m |v|^2 (v: the stored, half-step velocity)
     [1..3] momentum sum m vi, [4..6] angular momentum sum m x cross vi,
            where vi are the integer-step velocities of the last
            pmgpu_leapfrog(PMGPU_INTEGER_STEP), else v
     [7] sum over cells of rho * phi (the mesh potential energy is
         1/2 H^3 times this, after unit conversion)
     [8] total mass, [9..11] sum m x.
   Bitwise reproducible (fixed reduction order). */
#define PMGPU_DIAG_COUNT 12
int pmgpu_diagnostics(pmgpu_ctx* ctx, double* out);

/* ---- grids (Grid accessors, include/grid.h:18-44) -------------------- */
#define PMGPU_GRID_DENSITY 0   /* Nx*Ny*Nz floats, x fastest */
#define PMGPU_GRID_POTENTIAL 1 /* Nx*Ny*Nz floats */
#define PMGPU_GRID_FIELD 2     /* Nx*Ny*Nz*3 floats, (Ex,Ey,Ez) per cell */
#define PMGPU_GRID_GREENS 3    /* (Nx/2+1)*Ny*Nz floats: symmetrised G, times Nx*Ny*Nz */
int pmgpu_get_grid(pmgpu_ctx* ctx, int32_t which, float* out);

/* ---- learned influence-function correction --------------------------- */
/* Multiplies G(k) by (1 + MLP(kx/kN, ky/kN, kz/kN, order)); weights is the
   packed parameter vector of pmgpu_mlp_num_params() floats, NULL restores
   the analytic G.  Takes effect at the next pmgpu_init_greens. */
int pmgpu_set_greens_correction(pmgpu_ctx* ctx, const float* weights);
int64_t pmgpu_mlp_num_params(void);

/* ---- ensembles -------------------------------------------------------- */
/* No counterpart in the reference.  nsim independent simulations (each with
   its own pmgpu_pm_params, and pmgpu_p3m_params when use_p3m[i] != 0),
   stepped together by one HIP graph launch per two run-loop iterations.
   lanes (1..16) is the number of concurrent chains in the graph.  prepare
   computes the Green's functions, runs the initial force, half kick and
   first drift, autotunes the launch knobs per class of member when asked,
   and captures the graphs.  Steps must be even.  Members are periodic
   (escapes are recorded but do not stop a member). */
typedef struct pmgpu_batch pmgpu_batch;
int pmgpu_batch_create(int32_t nsim, const pmgpu_pm_params* params, const pmgpu_p3m_params* p3m,
                       const int32_t* use_p3m, int32_t lanes, pmgpu_batch** out);
void pmgpu_batch_destroy(pmgpu_batch* batch);
const char* pmgpu_batch_last_error(const pmgpu_batch* batch);
int pmgpu_batch_size(const pmgpu_batch* batch);
int pmgpu_batch_set_particles(pmgpu_batch* batch, int32_t sim, int64_t n, const float* pos, const float* vel,
                              const float* mass);
int pmgpu_batch_prepare(pmgpu_batch* batch, int32_t autotune);
int pmgpu_batch_step(pmgpu_batch* batch, int32_t steps);
int pmgpu_batch_synchronize(pmgpu_batch* batch);
/* wall milliseconds of `steps` ensemble steps (events on the ensemble stream) */
int pmgpu_batch_time(pmgpu_batch* batch, int32_t steps, double* ms);
/* per-member step latency (each member's own graph, alone), ms[nsim] */
int pmgpu_batch_member_latency(pmgpu_batch* batch, int32_t reps, double* ms);
int pmgpu_batch_get_particles(pmgpu_batch* batch, int32_t sim, float* pos, float* vel, float* acc);
int pmgpu_batch_get_tuning(pmgpu_batch* batch, int32_t sim, int32_t* knobs);

/* ---- launch tuning ------------------------------------------------------ */
/* No counterpart in the reference.  knobs[0] particles per deposit work item
   (1024..16384, default 1024), knobs[1] particles per interpolation work item
   (256..16384, default 2048), knobs[2] workgroups per CU of the
   interpolation kernel (1..32, default 32; the deposit launches four times
   as many), knobs[3] the
   short-range cell division: cells of y and z side >= cutoff / knobs[3]
   (x cells up to 4 times finer while the grid keeps within 128 cells per
   particle) and a
   (2 knobs[3] + 1)^2 neighbourhood of rows (1..6; unless set here,
   p3m_configure picks the largest from 3 whose y-z cell grid times the x
   cells at the same side has at most 16M cells (and at most 128 per
   particle once the particles are set), else 2, or 1 when even 2
   would give more than 128M cells, and with the particles set may then
   shape the cells to them: pmgpu_p3m_cell_geometry), knobs[4] the
   interpolation variant: 1 tiles staging the potential in LDS, 2 a gather
   from an E grid, 3 a gather straight from the potential (no E grid), 0
   automatic (the gather for grids of fewer than 4 tiles per CU, else the
   tiles);
   all give the same result up to rounding; knobs[5] the short-range work items: 1 chunks
   of one cell's particles, 2 chunks of a row of cells' particles (fewer idle
   lanes in sparse cells), 0 automatic (rows).  The PM results do not depend
   on the knobs (the deposit's fixed-point sums are order-independent); the
   short-range sum order, and so its rounding, follows knobs[3] and [5]. */
#define PMGPU_TUNING_KNOBS 6
int pmgpu_set_tuning(pmgpu_ctx* ctx, const int32_t* knobs);
int pmgpu_get_tuning(pmgpu_ctx* ctx, int32_t* knobs);

/* ---- profiling -------------------------------------------------------- */
/* on = 1: every kernel of the force step is bracketed by HIP events on the
   context's stream; on = 2: only the short-range pair kernel ("sr_pairs"),
   so that a timed run carries two event records per step; 0: off.  While
   either is on, pmgpu_run enqueues every iteration directly (no graph
   replay).  pmgpu_profile_read returns per-stage totals (milliseconds) and
   launch counts.  Names are static strings. */
int pmgpu_profile_enable(pmgpu_ctx* ctx, int32_t on);
int pmgpu_profile_reset(pmgpu_ctx* ctx);
int pmgpu_profile_read(pmgpu_ctx* ctx, int32_t max, const char** names, double* total_ms,
                       int64_t* launches, int32_t* count);
/* Ordered pairs with 0 < r < cutoff in the current state, and (tested, may be
   NULL) all candidate pairs the short-range pass visits.  Not on the step
   path: used to express the short-range kernel's rate in algorithmic
   FLOP/s. */
int pmgpu_p3m_count_pairs(pmgpu_ctx* ctx, int64_t* pairs, int64_t* tested);
/* stats[0] accepted pairs, stats[1] tested pairs, stats[2] lane-slots issued
   (64 per visited j, idle lanes of partly filled i chunks included). */
int pmgpu_p3m_pair_stats(pmgpu_ctx* ctx, int64_t* stats);
/* No counterpart in the reference.  The short-range cell geometry in use:
   geo[0..2] the x, y and z divisions (cell sides >= cutoff / geo[a]; the
   neighbourhood reaches geo[a] cells each way), geo[3] the number of cells.
   Unless knobs[3] of pmgpu_set_tuning fixes an isotropic division,
   pmgpu_p3m_configure (with particles set), and pmgpu_set_particles when it
   changes the particle count of a configured context, pick, among the isotropic
   division d and cells d times 4 or 2 finer in y or z and one cutoff (or
   half of one) deep in the other, the geometry whose candidate pairs plus
   16 per cell are fewest; all visit the same pair set. */
int pmgpu_p3m_cell_geometry(pmgpu_ctx* ctx, int32_t* geo);

/* ---- learned correction: standalone MLP evaluation ------------------- */
/* No counterpart in the reference.  inputs is n x 4 (host); out[m] = y(m).
   pmgpu_mlp_backward returns the gradient of sum_m up[m] * y(m) with respect
   to the packed parameters (BF16 MFMA, FP32 accumulation). */
int pmgpu_mlp_forward(const float* params, int64_t n, const float* inputs, float* out);
int pmgpu_mlp_backward(const float* params, int64_t n, const float* inputs, const float* up, float* grad);

/* ---- learned correction: training ------------------------------------ */
/* No counterpart in the reference; the loss is the reference's accuracy
   measure (script/pm_accuracy.py:44-48) as e_rms^2 = sum|a - t|^2 /
   sum|t|^2 of the PM force at target particles against given target
   accelerations.  A trainer owns a grid (params: grid, interpolation,
   finite_diff, greens_function, particle_diameter; box <= 0, periodic) and
   a set of snapshots: positions and masses in code units, target indices
   into them and the targets' accelerations (code units).  The correction
   multiplies the trainer's analytic G by (1 + y(k)). */
typedef struct pmgpu_trainer pmgpu_trainer;
int pmgpu_trainer_create(const pmgpu_pm_params* params, pmgpu_trainer** out);
void pmgpu_trainer_destroy(pmgpu_trainer* trainer);
const char* pmgpu_trainer_last_error(const pmgpu_trainer* trainer);
int pmgpu_trainer_add_snapshot(pmgpu_trainer* trainer, int64_t n, const float* pos, const float* mass,
                               int64_t n_targets, const int64_t* target_index, const float* target_acc);
/* Loss for params (NULL = analytic G); grad (NULL to skip) receives dL/dparams,
   grad_c (NULL to skip) dL/d(1 + y(k)) on the half spectrum [z][y][kx]. */
int pmgpu_trainer_loss_grad(pmgpu_trainer* trainer, const float* params, double* loss, float* grad,
                            float* grad_c);
/* Adam (beta 0.9 / 0.999) for `steps` steps on the joint loss of all
   trainers (one normaliser over all their targets); params is updated in
   place; loss_history (may be NULL) receives the loss before each step. */
int pmgpu_train(pmgpu_trainer** trainers, int32_t n_trainers, float* params, int32_t steps, float lr,
                float* loss_history);

/* ---- direct-sum forces (training targets and error checks) ------------ */
/* a_i = -1/(4 pi) sum_j m_j d_ij / (|d_ij|^2 + eps^2)^(3/2), d_ij = x_i - x_j
   (code units, Plummer softening as source/ppMethod.cpp:23-36 with G' =
   1/(4 pi)), coincident pairs skipped; FP32 with compensated sums.  Host
   arrays: pos n x 3, mass n, tpos nt x 3, out nt x 3.  box NULL: isolated;
   else the periodic force in a box of box[0..2] (nearest image plus a
   tabulated Ewald correction on a neutralising background, the boundary
   condition of the FFT force step). */
int pmgpu_direct_sum(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                     float softening, const float* box, float* out);
/* pmgpu_direct_sum without any device allocation (no counterpart in the
   reference): for callers that cannot allocate device memory -- serving
   loops, harnesses that forbid hipMalloc -- and for loops that would
   otherwise allocate per call.  The staging copies of the inputs and
   outputs and the Ewald table live in one static device array of about
   40 MiB per device, held until the process ends; at most
   PMGPU_DS_NOALLOC_MAX_SOURCES sources and PMGPU_DS_NOALLOC_MAX_TARGETS
   targets (PMGPU_ERR_UNSUPPORTED beyond).  Same arguments, kernels and
   results as pmgpu_direct_sum; runs on the null stream of the current
   device and returns when out is written.  Calls that overlap on one
   device share that array and race. */
#define PMGPU_DS_NOALLOC_MAX_SOURCES (1 << 20)
#define PMGPU_DS_NOALLOC_MAX_TARGETS (1 << 16)
int pmgpu_direct_sum_noalloc(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                             float softening, const float* box, float* out);
/* The same sums on the host CPU in FP64 over `threads` threads (double
   accumulation; the periodic correction from the same Ewald table), for
   error statements against a direct sum.  out: nt*3 doubles. */
int pmgpu_direct_sum_host(int64_t n, const float* pos, const float* mass, int64_t nt, const float* tpos,
                          float softening, const float* box, int32_t threads, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PMGPU_H */
