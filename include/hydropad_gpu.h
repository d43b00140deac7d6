/*
 * hydropad_gpu.h -- C-ABI of the MI355X-native directionally split PPM hydro step.
 *
 * This is the drop-in boundary for Hydropad's hydro hot path.  Every entry
 * point names the reference interface it replaces (paths relative to the
 * cgheller/hydropad tree).  All arguments are plain pointers, sizes and
 * scalars; no C++ or framework types cross this boundary, so the header can be
 * bound from Fortran (ISO_C_BINDING, see hydropad_amd/fortran/hydropad_gpu.F90),
 * from C, or from Python ctypes.
 *
 * Conventions
 *  - Fields are REAL(8) arrays in Fortran (column-major, x fastest) order with
 *    dimensions (nx,ny,nz) = (N+2*nbound, ...), ghosts included, exactly as the
 *    reference allocates p3d/rho3d/vx3d/vy3d/vz3d (hydropad/alloc_arrays.F90:30-32,56-60).
 *  - gforce, when given, is REAL(8) (3,nx,ny,nz) interleaved (hydropad/alloc_arrays.F90:89).
 *  - Every call returns 0 on success and a negative HP_E* code on failure; the
 *    reference's `stop` on Newton non-convergence (hydro/riemann.F90:60-65)
 *    becomes a device counter, and hp_ppmsolver, hp_ppm_lxlylz, hp_evolve and
 *    hp_evolve_cosmo return HP_E_NEWTON when it is non-zero (it counts since
 *    the last hp_reset_counters; the state past such a step is what the
 *    reference would never have produced).
 *  - A context owns its device state and one HIP stream; calls on one context
 *    are stream-ordered and must come from one host thread at a time.
 */
#ifndef HYDROPAD_GPU_H
#define HYDROPAD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define HP_OK            0
#define HP_E_ARG        -1   /* bad argument (null pointer, bad size, bad direction) */
#define HP_E_HIP        -2   /* a HIP runtime call failed */
#define HP_E_NOMEM      -3   /* device allocation failed / budget exceeded */
#define HP_E_NEWTON     -4   /* Riemann Newton iteration exceeded 100 steps (hydro/riemann.F90:60-65) */
#define HP_E_STATE      -5   /* call not valid in the context's current state */

/* ---- precision forms ---------------------------------------------------- */
#define HP_FP64          0   /* everything in FP64 (throughput arithmetic, same algorithm) */
#define HP_MIXED         1   /* FP32 reconstruction/tracing/fluxes; FP64 Newton + dual energy + state */
#define HP_FP64_EXACT    2   /* FP64 in the reference's own operation order, no FMA contraction:
                                bit-faithful to the reference up to OCML-vs-glibc pow rounding */

/*
 * The 19 module scalars the reference pushes to the device before every sweep
 * (hydropad/scalar.F90:98-99, hydro/ppmsolver.F90:68-69).  The gamma-derived
 * members follow hydropad/initialize_parameters.F90:64-69, dx/rdx follow
 * hydropad/initialize_parameters.F90:23-24 (dx is rounded through FP32 there).
 */
typedef struct hp_params {
    double gamma, gamma1, rgamma, rgamma1, m, rm, gf;
    double dt, dx, rdx, rdtath, at, dat, rat;
    double flatvalue, mingradflat, eta1, eta2, dmax;
} hp_params;

/* Counters that replace the reference's prints and `stop`s
 * (hydro/order2interp.F90:174-190, hydro/riemann.F90:62-65, hydro/integrate.F90:190-217). */
typedef struct hp_counters {
    int64_t newton_iters_max;   /* largest Newton iteration count seen */
    int64_t newton_nonconv;     /* interfaces whose Newton loop hit 100 iterations */
    int64_t neg_state_clamps;   /* traced p/rho < 0 clamped to 1e-15 */
    int64_t particles_outside;  /* kicks skipped: the particle's CIC cube leaves the array */
    int64_t reserved[4];
} hp_counters;

typedef struct hp_ctx hp_ctx;

/* Fill the derived constants exactly as hydropad/initialize_parameters.F90:23-24,64-69
 * and :183-203 (non-cosmological time init: at=1, dat=0, rat=1, rdtath=dt).
 * ngrid_max is max(ngridx,ngridy,ngridz).  Host-only; never touches a GPU. */
int hp_params_init(double gamma, int ngrid_max, double dt, hp_params* out);

/* ---- context ------------------------------------------------------------ */
/* Replaces the OpenACC `declare create` of the module fields (hydropad/matrix.F90:24)
 * and the per-sweep temporaries of hydro/ppmsolver.F90:55-64: device state is
 * allocated once here and stays resident. precision is HP_FP64, HP_MIXED or HP_FP64_EXACT. */
int  hp_create(int nx, int ny, int nz, int nbound, int precision, hp_ctx** out);
void hp_destroy(hp_ctx* ctx);
int  hp_set_precision(hp_ctx* ctx, int precision);

/* Host<->device copies of the 5 fields (ghosts included) and the optional
 * gravity array; replaces `!$acc update device/host` (hydro/ppmsolver.F90:95,
 * hydro/ppm_lxlylz.F90:20).  gforce may be NULL (no gravity, g == 0). */
int hp_upload(hp_ctx* ctx, const double* p, const double* rho, const double* vx,
              const double* vy, const double* vz, const double* gforce);
int hp_download(hp_ctx* ctx, double* p, double* rho, double* vx, double* vy, double* vz);

/* The gravity array alone, host (3,nx,ny,nz) REAL(8), or NULL for none; the
 * fields are left as they are.  The reference sets gforce once
 * (hydropad/initialize_gravity.F90:25-58) and both the sweeps
 * (hydro/ppmsolver.F90:108,151) and kick (nbody/kick.F90:1) read it. */
int hp_upload_gforce(hp_ctx* ctx, const double* gforce);

/* Device-to-device variants: pointers are device (or managed) memory. */
int hp_upload_device(hp_ctx* ctx, const double* p, const double* rho, const double* vx,
                     const double* vy, const double* vz, const double* gforce);
int hp_download_device(hp_ctx* ctx, double* p, double* rho, double* vx, double* vy, double* vz);

/* The module scalars for the next sweep(s); replaces hydro/ppmsolver.F90:68-69. */
int hp_set_params(hp_ctx* ctx, const hp_params* prm);

/* ---- the hot path --------------------------------------------------------- */
/* One directional sweep in place on the device state, ndirection 1=x 2=y 3=z.
 * Replaces SUBROUTINE ppmsolver (hydro/ppmsolver.F90:1-467, STENCIL branch
 * :94-160, :213-288, :341-416).  On return the face ghosts of all three
 * directions are periodic-consistent, as after the reference's exchange_mesh. */
int hp_ppmsolver(hp_ctx* ctx, int ndirection);

/* Periodic ghost fill of the 5 fields on the device; replaces SUBROUTINE
 * exchange_mesh (hydropad/exchange_mesh.F90:2-126, exchange_var.F90:2-57) on one
 * rank.  direction 1..3 fills that axis' faces; 0 fills all, edges and corners
 * included, in the order z, x, y of exchange_mesh.F90:47-123. */
int hp_exchange_mesh(hp_ctx* ctx, int direction);

/* One full hydro step: three sweeps in the order given by nrot (1..6) with the
 * ghost fills in between; replaces SUBROUTINE ppm_lxlylz (hydro/ppm_lxlylz.F90:1-223). */
int hp_ppm_lxlylz(hp_ctx* ctx, int nrot);

/* CFL maximum velocity max(|v_d| + c_s) over interior cells, reduced on the
 * device; replaces SUBROUTINE speed (hydropad/speed.F90:2-59) incl. its
 * MPI_Allreduce.  Only 8 bytes cross PCIe. */
int hp_speed(hp_ctx* ctx, double* velmax);

/* Device-resident time loop shaped like hydropad/main.F90:100-121 +
 * evolve_sys.F90:30-31,54,122-129: for nstep = first_step .. first_step+nsteps-1,
 * nrot = mod(nstep,6)+1 (1 on step 1), ppm_lxlylz, speed, timedet
 * (hydropad/functions.F90:2-35).  cour is the Courant number; dt/t are in/out.
 * No host synchronisation happens inside the loop. */
int hp_evolve(hp_ctx* ctx, int first_step, int nsteps, double cour, double tstop,
              double* t, double* dt);

/* ---- comoving (cosmological) boxes ----------------------------------------
 * Einstein-de Sitter schedule state, the members of hydropad/scalar.F90:16-19
 * that timedet_cosmo reads and writes.  at/dat are a and da/dt at the start
 * of the upcoming step, ath/dath at its midpoint, atnew/datnew at its end,
 * t its end time and dt its length; nstep counts completed steps. */
typedef struct hp_cosmo {
    double cour, ca, tstop;      /* Courant number, max da/a per step, stop time */
    double t, dt;
    double at, dat, ath, dath, atnew, datnew;
    double rat, rdtath;          /* 1/at, dt/ath */
    int64_t nstep;
} hp_cosmo;

/* Initial schedule for a box starting at expansion factor a_start; replaces
 * the COSMO / omega_m = 1 branch of hydropad/initialize_parameters.F90:136-178
 * (tin = a^1.5, first dt = 0.010*tin).  Host-only. */
int hp_cosmo_init(double a_start, double cour, double ca, double tstop, hp_cosmo* s);

/* SUBROUTINE timedet_cosmo, omega_m = 1 branch (hydropad/functions.F90:38-151),
 * after step s->nstep+1 with that step's CFL maximum velmax; dx is the cell
 * size.  The reference's dt0 = 1e-9 cap (functions.F90:49,97) is not applied
 * and there are no particle speeds.  Host-only. */
int hp_timedet_cosmo(hp_cosmo* s, double velmax, double dx);

/* Copy the schedule's at, dat, rat, dt, rdtath into the sweep scalars. */
int hp_cosmo_params(const hp_cosmo* s, hp_params* prm);

/* Keep p, vx, vy, vz of the current state as the start-of-step state the
 * expansion step uses (the reference's vxold/vyold/vzold/pold3d, hydropad/matrix.F90). */
int hp_savetn(hp_ctx* ctx);

/* SUBROUTINE expand (hydropad/expand.F90:1-43) on the device, over the whole
 * array, with afact = dt*(dath/ath) from the caller. */
int hp_expand(hp_ctx* ctx, double afact);

/* Device-resident cosmological time loop: for each step, ppm_lxlylz in the
 * nrot order with the schedule's scalars, expand (fused into the step's last
 * sweep), speed and timedet_cosmo (on the device) -- hydropad/evolve_sys.F90
 * with COSMO defined.  The schedule *s is read at entry and written back at
 * return. */
int hp_evolve_cosmo(hp_ctx* ctx, int first_step, int nsteps, hp_cosmo* s);

/* Launch shape of this context's sweeps: x_chunks = 58-cell chunks each
 * x-sweep wave walks (0: about 16 resident rounds of waves), yz_seglen =
 * y/z pencil segment length (0: at least 64 cells, short enough for ~16K
 * waves), block = threads per block
 * (0: 256; a multiple of 64, at most 256).  The shape never changes a result.
 * No reference counterpart (the reference's OpenACC gang/vector shapes are
 * compiler-chosen). */
int hp_set_tiling(hp_ctx* ctx, int x_chunks, int yz_seglen, int block);

/* Per-sweep kernel timing with HIP events on the context's stream; replaces
 * the MPI_WTIME "Elapsed time" print around every ppmsolver call
 * (hydro/ppmsolver.F90:72,464-465).  While enabled, every sweep launch is
 * bracketed by two events; hp_sweep_timing synchronises and returns, per
 * direction (index 0..2 = x,y,z), the summed kernel milliseconds and the
 * number of launches since the last enable. */
int hp_sweep_timing_enable(hp_ctx* ctx, int on);
int hp_sweep_timing(hp_ctx* ctx, double ms[3], int64_t launches[3]);

/* ---- particles (the N-body step beside the hydro step) -----------------------
 * The reference's ppos/pvel, (3,np) REAL(8) interleaved (nbody/kick.F90:16-17),
 * in the units of hydropad/init_sys.F90:93-95 (the box is [0,boxsize) with
 * boxsize = ngrid/max(ngrid) in FP32).  Particles with ppos(1) == -1 are
 * inactive (kick.F90:37).  dtat is the kicks' time step, the reference's
 * module scalar dtat (hydropad/scalar.F90:18), which it sets once at
 * initialisation (initialize_parameters.F90:174,193) and never updates; the
 * time loops below keep it fixed in the same way.  When particles are present,
 * hp_evolve and hp_evolve_cosmo run kick, drift, kick after the sweeps of each
 * step (hydropad/evolve_sys.F90:63-84, one rank) with the context's gforce
 * (zero without one), and hp_speed and the loops' CFL maximum include max |v|
 * of every particle (hydropad/speed.F90:43-56). */
int hp_particles_upload(hp_ctx* ctx, int64_t np, const double* ppos, const double* pvel, double dtat);
int hp_particles_download(hp_ctx* ctx, double* ppos, double* pvel);

/* SUBROUTINE kick (nbody/kick.F90:1-133): v <- v (1 - dat dtat/2) + F dtat/2,
 * F the cell-in-cell interpolation of the context's gforce (the sweeps'
 * scalars supply dat).  A particle whose interpolation cube would leave the
 * array is not kicked and is counted in hp_counters.particles_outside. */
int hp_kick(hp_ctx* ctx, double dtat);

/* SUBROUTINE drift (nbody/drift.F90:1-36): x <- x + rdtath v, wrapped once
 * into [0, boxsize). */
int hp_drift(hp_ctx* ctx);

/* Accumulated counters since the last hp_reset_counters; returns HP_E_NEWTON
 * when newton_nonconv is non-zero. */
int hp_status(hp_ctx* ctx, hp_counters* out);
int hp_reset_counters(hp_ctx* ctx);

/* Iteration count past which a Riemann Newton iteration counts as not
 * converged: the reference's local nnewtmax = 100 (hydro/riemann.F90:51,60-65).
 * nmax = 0 restores 100.  A smaller limit makes the non-convergence path
 * reachable in tests (finite positive states converge well within 100). */
int hp_set_newton_limit(hp_ctx* ctx, int nmax);
int hp_sync(hp_ctx* ctx);

/* ---- batched ensemble ----------------------------------------------------------
 * Many independent periodic boxes, each with its own size, precision, step
 * count, time step and (for comoving boxes) expansion schedule, resident on
 * one GPU and advanced together.  Each batched step runs, for every box, one
 * step of the evolve_sys loop above (its own nrot order, hydropad/evolve_sys.F90:30-31;
 * its own timedet or timedet_cosmo): a plan kernel, three sweep slots of
 * persistent per-direction kernels, one timedet kernel -- no host work and
 * no synchronisation, so the sequence is captured in a hipGraph when asked.
 * The reference has no counterpart (it advances one box per MPI job); the
 * per-box arithmetic is that of HP_FP64 / HP_MIXED, and the tiling does not
 * change any result.  Gravity is not supported in a batch. */
typedef struct hp_batch hp_batch;

typedef struct hp_batch_tuning {
    int x_chunks;         /* x sweep: 58-cell chunks per work unit (0: 4) */
    int yz_segments;      /* y/z sweeps: segments each pencil is split into (0: 1) */
    int waves_per_cu;     /* persistent grid: waves per compute unit (0: 8) */
    int reserved;
} hp_batch_tuning;

typedef struct hp_box_status {
    int64_t steps_done;
    int32_t active, done;
    double t, dt;         /* end time of the step in flight / next step, its dt */
    int64_t newton_iters_max, newton_nonconv, neg_state_clamps;
} hp_box_status;

/* Device bytes of one box's field sets ((nx,ny,nz) with ghosts; comoving
 * boxes keep a third set for the expansion step). */
size_t hp_batch_box_bytes(int nx, int ny, int nz, int comoving);

int  hp_batch_create(int max_boxes, size_t hbm_budget_bytes, hp_batch** out);
void hp_batch_destroy(hp_batch* b);

/* Add a box from host arrays (Fortran order, ghosts included) to run nsteps
 * steps.  cosmo == NULL: static box, first dt from its own CFL condition
 * (speed + timedet with t = 0, cour, tstop); otherwise a comoving box on that
 * schedule.  Returns HP_E_NOMEM when the slot table or the HBM budget is full. */
int hp_batch_add(hp_batch* b, int nx, int ny, int nz, int nbound, int precision, double gamma,
                 double cour, double tstop, const hp_cosmo* cosmo, int nsteps,
                 const double* p, const double* rho, const double* vx, const double* vy,
                 const double* vz, int* id);

/* Add an n^3 static box holding the seeded planar Sod tube, generated on the
 * device (gamma = 5/3, no upload); see hp_sod_of_seed. */
int hp_batch_add_sod(hp_batch* b, int n, int precision, uint64_t seed, double cour, int nsteps, int* id);

/* What a seed selects for the Sod tube of hp_batch_add_sod: the axis (0..2),
 * the split position x0 in [0.25, 0.75) and whether the high-pressure side is
 * flipped to x >= x0 (from splitmix64 of the seed).  Host-only. */
int hp_sod_of_seed(uint64_t seed, int* axis, double* x0, int* flip);

int hp_batch_remove(hp_batch* b, int id);

/* hp_set_newton_limit for one box of the batch (0 restores the reference's 100). */
int hp_batch_set_newton_limit(hp_batch* b, int id, int nmax);
int hp_batch_download(hp_batch* b, int id, double* p, double* rho, double* vx, double* vy, double* vz);

/* Enqueue nsteps batched steps (asynchronous); boxes that reached their step
 * count are left alone. */
int hp_batch_step(hp_batch* b, int nsteps);
int hp_batch_sync(hp_batch* b);

/* Synchronise and report every slot (array of max_boxes entries). */
int hp_batch_poll(hp_batch* b, hp_box_status* out, int* nactive);

/* Capture each batched step in a hipGraph and replay it (re-captured when the
 * mix of precisions / comoving boxes or the grid changes). */
int hp_batch_use_graph(hp_batch* b, int on);

int hp_batch_set_tuning(hp_batch* b, const hp_batch_tuning* t);
int hp_batch_get_tuning(hp_batch* b, hp_batch_tuning* t);
size_t hp_batch_bytes_used(const hp_batch* b);

/* Autotuner: coordinate search over x_chunks, yz_segments and waves_per_cu,
 * each trial steps_per_trial real batched steps timed with HIP events, the
 * candidates of each parameter timed in order and in reverse order and
 * scored by their mean (22 trials in all); leaves the fastest tiling set.
 * *steps_run receives the steps it advanced. */
int hp_batch_autotune(hp_batch* b, int steps_per_trial, hp_batch_tuning* best,
                      double* best_ms_per_step, int* steps_run);

/* ---- serving ---------------------------------------------------------------
 * A request loop with continuous batching: requests (a box size, a step
 * count, a precision and a seed for the Sod initial state) arrive at
 * arrival_s seconds after the call starts, are admitted in arrival order
 * while box slots and the HBM budget allow, advance one batched step at a
 * time together with everything resident, and retire when their step count
 * is reached.  result_fields, if not NULL, holds 5 host pointers per request
 * (p, rho, vx, vy, vz of (n+6)^3 doubles, or NULL) for the final states. */
typedef struct hp_request {
    int64_t id;
    int32_t n, nsteps;
    int32_t precision;
    int32_t newton_limit;   /* 0: the reference's 100 (see hp_set_newton_limit) */
    uint64_t seed;
    double arrival_s;
    double cour;
} hp_request;

typedef struct hp_request_result {
    int64_t id;
    double arrival_s, admit_s, done_s;
    int64_t steps;
    double t;
    int64_t newton_nonconv;
    int32_t rc, pad;
} hp_request_result;

typedef struct hp_serve_stats {
    double wall_s, cell_updates;
    double p50_s, p99_s, mean_s;
    int64_t batched_steps, max_concurrent, completed;
    double peak_bytes;
} hp_serve_stats;

int hp_serve_run(const hp_request* reqs, int nreq, size_t hbm_budget, int max_boxes, int use_graph,
                 double* const* result_fields, hp_request_result* results, hp_serve_stats* stats);

/* ---- one sweep with no device allocation -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hp_ppmsolver's sweep (same kernels, same results) on
 * caller-owned device buffers.  p..vz are the input fields and p_out..vz_out
 * the output fields, (nx,ny,nz) REAL(8) device arrays in the reference layout
 * (ghosts included; the output's face ghosts are written, edges and corners
 * are not); input and output must not overlap.  gforce is a device
 * (3,nx,ny,nz) array or NULL.  newton_limit as hp_set_newton_limit (0: 100).
 * What hp_create would allocate besides the fields -- the device scalars and
 * the status block -- is a library-owned array of HP_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends and shared by every call on that GPU:
 * calls that overlap on one GPU race (one stream at a time).  The call is
 * stream-ordered on `stream` (a hipStream_t; NULL = the default stream) and
 * calls no allocating routine.  counters == NULL: fully asynchronous;
 * otherwise it synchronises the stream, fills *counters and returns
 * HP_E_NEWTON on Newton non-convergence. */
#define HP_NOALLOC_SCRATCH_BYTES 512
int hp_ppmsolver_noalloc(const hp_params* prm, const double* p, const double* rho, const double* vx,
                         const double* vy, const double* vz, double* p_out, double* rho_out, double* vx_out,
                         double* vy_out, double* vz_out, const double* gforce, int nx, int ny, int nz,
                         int nbound, int ndirection, int precision, int newton_limit, hp_counters* counters,
                         void* stream);

/* ---- scalar entry points (host-callable, for unit vectors) ---------------- */
/* SUBROUTINE integrate (hydro/integrate.F90:1-113) on one cell, arguments by
 * pointer in the reference's order; pres, rho, vx, vy, vz are updated in place.
 * The caller's etot and eint are used as given (the internal-energy equation
 * and its fallbacks, integrate.F90:75-81,102-105, start from eint).
 * pmean, undefined in the reference's active build, is (pm1+pm2)*0.5 as in
 * hydro/ppm.F90:583. */
int hp_integrate_cell(const hp_params* prm,
                      const double* pm1, const double* pm2, const double* rhom1, const double* rhom2,
                      const double* vxm1, const double* vxm2, const double* vym1, const double* vym2,
                      const double* vzm1, const double* vzm2,
                      double* pres, double* rho, double* vx, double* vy, double* vz,
                      const double* etot, const double* eint, const double* ghalf);

/* The dual-energy selection inlined at hydro/integrate.F90:182-217:
 * given the updated total energy etotnew, density and velocities, the
 * internal-energy-equation value eintnew_in and the old internal energy and
 * pressure, returns the new pressure and internal energy. */
int hp_dualenergy_cell(const hp_params* prm, double etotnew, double rhonew, double vx,
                       double vy, double vz, double eintnew_in, double eint_old,
                       double pres_old, double* pnew, double* eintnew);

#ifdef __cplusplus
}
#endif
#endif /* HYDROPAD_GPU_H */
