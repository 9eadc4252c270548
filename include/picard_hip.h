/*
 * picard_hip.h -- C ABI of the MI355X (gfx950) implementation of Picard's
 * PIC time step.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status (PICARD_OK = 0, negative = error, positive = warning);
 * picard_last_error() gives the message of the last failure on the calling
 * thread.  State (particles, fields) stays resident on the device inside a
 * picard_ctx between calls; nothing is copied to or from the host unless a
 * *_read / *_write / *_append entry point is called.
 *
 * The Fortran drop-in layer (picard_amd/fortran/picard_dropin.f90) binds these
 * through ISO_C_BINDING and exposes the reference's subroutine names and
 * argument lists on top of them.
 *
 * Array layouts follow the reference (Fortran column-major, one guard layer):
 *   real_particles(8, n)  -> double[8*n], slot fastest
 *   int_particles(n)      -> int32_t[n]  (species, 1-based)
 *   U, C (Nx+2,Ny+2,Nz+2) -> double[G],  G = (Nx+2)(Ny+2)(Nz+2)
 *   E(3, ...)             -> double[3*G], component fastest
 *   N(species, ...)       -> double[S*G], species fastest
 *   D(2Nx+1,2Ny+1,2Nz+1)  -> double[(2Nx+1)(2Ny+1)(2Nz+1)]
 */
#ifndef PICARD_HIP_H
#define PICARD_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PICARD_OK = 0,
  PICARD_ERR_ARG = -1,       /* bad argument or shape */
  PICARD_ERR_HIP = -2,       /* HIP runtime failure */
  PICARD_ERR_CAPACITY = -3,  /* more particles than max_particles */
  PICARD_ERR_RANGE = -4,     /* a particle's CIC stencil falls outside the grid arrays */
  PICARD_ERR_STATE = -5,     /* call out of order (e.g. BCpotentials before CalcDistance) */
  PICARD_WARN_NOCONV = 1     /* CalcEpotential ran past maxit without converging */
};

enum { PICARD_FP64 = 0, PICARD_FP32_SPLIT = 1 };
enum { PICARD_DEPOSIT_REFERENCE_ORDER = 0, PICARD_DEPOSIT_FIXED_POINT = 1 };
enum {
  PICARD_FIELD_U = 0, PICARD_FIELD_E = 1, PICARD_FIELD_N = 2, PICARD_FIELD_C = 3,
  PICARD_FIELD_D = 4,
  PICARD_FIELD_F = 5   /* number fluxes F(species, 3, nodes), after particles_gridify_all */
};

/* One domain (one MPI rank of the reference, here one GPU). */
typedef struct {
  int32_t nx, ny, nz;        /* Nx_local, Ny_local, Nz_local (picard.f90:46-48) */
  int32_t species;           /* species (picard.f90:50), 1..8 */
  int32_t flow[3];           /* xflow, yflow, zflow (picard.f90:148-150), nonzero = .true. */
  int32_t precision;         /* PICARD_FP64 or PICARD_FP32_SPLIT */
  int32_t deposit;           /* PICARD_DEPOSIT_* (FP64 form; FP32-split always fixed point) */
  int32_t device;            /* HIP device ordinal */
  int64_t max_particles;     /* max_per_proc (picard.f90:52) */
  double lmin[3], lmax[3];   /* Lx_min..Lz_max (picard.f90:350-355) */
  double gmin[3], gmax[3];   /* xmin..zmax (picard.f90:159-164) */
  double dxyz[3];            /* cell sizes (picard.f90:358-360) */
} picard_geometry;

/* Per-species constants derived by the host (picard.f90:365-466). */
typedef struct {
  double charge[8], n2p[8], eta[8];
  double tB[32], sB[32];     /* tB(4,8), sB(4,8): element (c,s) at [4*(s-1) + (c-1)] */
  double ve[4], B0[4], E0[4];
  int32_t fluid[8], gyration[8];
} picard_species;

/* Plasma composition used by the device-side sources (picard.f90:371-401). */
typedef struct {
  int32_t ipp[8], epp[8], hpp[8];   /* ions per cell; electrons, protons per ion */
  double mass[8], kelvin[8];
  double v0[32];                    /* v0(4,8): element (c,s) at [4*(s-1) + (c-1)] */
} picard_plasma;

typedef struct picard_ctx picard_ctx;

const char *picard_last_error(void);

/* Allocate a device-resident domain.  No reference counterpart: the
 * reference keeps these arrays as static locals of the main program
 * (picard.f90:91-113). */
int picard_create(const picard_geometry *geo, const picard_species *sp, picard_ctx **out);
int picard_destroy(picard_ctx *ctx);
int picard_set_species(picard_ctx *ctx, const picard_species *sp);
int picard_synchronize(picard_ctx *ctx);

/* Append `count` particles after the current num_local: the range that
 * BCinflow (BCinflow.f90:2) and ParticlesProduce (ParticlesProduce.f90:1)
 * write on the host.  real_particles points at the first appended particle,
 * 8 doubles each. */
int picard_particles_append(picard_ctx *ctx, const double *real_particles,
                            const int32_t *int_particles, int64_t count);
/* Copy particles [first, first+count) (0-based) back to host AoS arrays. */
int picard_particles_read(picard_ctx *ctx, int64_t first, int64_t count, double *real_particles,
                          int32_t *int_particles);
int picard_particles_count(picard_ctx *ctx, int64_t *num_local);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc):
 * picard_particles_read staged through a fixed 16 MiB device array, one per
 * GPU, instead of a staging buffer the library allocates.  At most
 * PICARD_READ_NOALLOC_MAX particles per call (larger counts return
 * PICARD_ERR_ARG; read in pieces with `first`); sp may be NULL.  Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU race. */
#define PICARD_READ_NOALLOC_MAX 246723
int picard_particles_read_noalloc(picard_ctx *ctx, int64_t first, int64_t count, double *rp,
                                  int32_t *sp);

/* Device-side particle sources, appended after num_local.  They replace the
 * host loops of ParticlesInitialize (ParticlesInitialize.f90:2) and BCinflow
 * (BCinflow.f90:2) with the same cell ranges, loop order and distributions,
 * drawing from a counter-based generator (Philox4x32-10) keyed by `seed`
 * instead of the Fortran runtime's random_number. */
int picard_generate_initial(picard_ctx *ctx, const picard_plasma *plasma, uint64_t seed);
int picard_generate_inflow(picard_ctx *ctx, const picard_plasma *plasma, uint64_t seed,
                           int64_t step);

/* Haser-model production parameters (picard.f90:258-302 set them; species
 * index s-1 holds species s).  rate_scale multiplies the expected ion count
 * per cell of ParticlesProduce.f90:70: 1.0 evaluates the line as written (a
 * rate per second taken as a count per step), dt makes it a per-step count. */
typedef struct {
  double Qn[8], vn[8], nu_i[8], nu_d[8], Ek[8];
  double rate_scale;
} picard_haser;

/* ParticlesProduce (ParticlesProduce.f90:1-113) on the device: the same
 * cells, loop order, counts and companion electrons/protons, drawn from the
 * counter-based generator keyed by `seed` and counted by `step`.  Appends
 * after num_local; *produced receives the number of particles appended. */
int picard_particles_produce(picard_ctx *ctx, const picard_plasma *plasma,
                             const picard_haser *haser, uint64_t seed, int64_t step, double dt,
                             int64_t *produced);
/* Make picard_step run picard_particles_produce after AdvancePosition, where
 * the reference calls ParticlesProduce (picard.f90:566), with step = 1, 2, ...
 * counted by picard_step.  haser == NULL switches the source off. */
int picard_set_production(picard_ctx *ctx, const picard_plasma *plasma,
                          const picard_haser *haser, uint64_t seed);

/* The diagnostics line of the main loop (picard.f90:649-672), computed on
 * the device: uP = 0.5 sum n2p(s) mass(s) |v|^2 over the particles (:658-661),
 * uE = 0.5 eps0 dV sum |E + E0|^2 over interior nodes (:655-657), uB = 0.5/mu0
 * dV sum |B0|^2 over interior nodes (B = B0 everywhere, :420-422, :653).  The
 * reference prints them divided by its initial energy u0 (:476-480). */
typedef struct {
  int64_t num_local, num_global;
  double particle_energy, e_energy, b_energy;
} picard_diag;

int picard_diagnostics(picard_ctx *ctx, const double *mass /* [8] */, picard_diag *out);

/* ---- the time-step call surface (picard.f90:564-631) ---- */

/* AdvancePosition(real_particles, dt, max_per_proc, num_local)   AdvancePosition.f90:4 */
int picard_advance_position(picard_ctx *ctx, double dt);

/* BCoutflow(real_particles, int_particles, Lx_min..Lz_max, xmin..zmax, max_per_proc,
 *           num_local, xflow, yflow, zflow)                        BCoutflow.f90:5
 * Reference swap-with-last order; num_local (out) and removed (out, optional). */
int picard_bc_outflow(picard_ctx *ctx, int64_t *num_local, int64_t *removed);

/* ParticlesTransfer(...) on one domain: periodic wrap     ParticlesTransfer.f90:2 */
int picard_particles_transfer(picard_ctx *ctx);

/* ParticlesGridify(N, C, real_particles, int_particles, ...)   ParticlesGridify.f90:1 */
int picard_particles_gridify(picard_ctx *ctx);
/* ParticlesGridifyAll (ParticlesGridifyAll.f90:1-455): ParticlesGridify's N
 * and C plus the number fluxes F(s, c, node) = sum n2p(s) W v_c with the same
 * guard folds; used on field-output steps (picard.f90:610-612).  FP64 form:
 * reference summation order; split form: fixed point. */
int picard_particles_gridify_all(picard_ctx *ctx);

/* CalcDistance(D, Nx, Ny, Nz, dxyz, xflow, yflow, zflow)        CalcDistance.f90:1 */
int picard_calc_distance(picard_ctx *ctx);

/* BCpotentials(U, C, D, xmin..zmax, Nx_local.., Nx.., flows, ...)  BCpotentials.f90:1 */
int picard_bc_potentials(picard_ctx *ctx);

/* CalcEpotential(U, C, ..., dxyz, maxerr, maxit, ...)        CalcEpotential.f90:4
 * sweeps (out) = Jacobi sweeps performed.  Returns PICARD_WARN_NOCONV when the
 * loop ran to 2*maxit+1 sweeps. */
int picard_calc_epotential(picard_ctx *ctx, double maxerr, int32_t maxit, int64_t *sweeps);

/* CalcEfield(E, U, Nx_local.., dxyz, flows, ...)                 CalcEfield.f90:3 */
int picard_calc_efield(picard_ctx *ctx);

/* BCfields(E, U, B0, E0, Nx_local.., flows, ...)                 BCfields.f90:3 */
int picard_bc_fields(picard_ctx *ctx);

/* AdvanceVelocity(real_particles, int_particles, E, B0, Lx_min.., dxyz, dt, v0, ve,
 *                 eta, tB, sB, num_local, ...)                   AdvanceVelocity.f90:6 */
int picard_advance_velocity(picard_ctx *ctx, double dt);
/* The output-step pair that replaces AdvanceVelocity on steps that write
 * particles or fields (picard.f90:602-621): AdvanceVelocityImp
 * (AdvanceVelocity.f90:240-347) takes v to the half step before the dumps,
 * AdvanceVelocityExp (:128-238) completes the step after them.  The main
 * loop passes deltat = dt/2. */
int picard_advance_velocity_imp(picard_ctx *ctx, double deltat, double dt);
int picard_advance_velocity_exp(picard_ctx *ctx, double deltat, double dt);

/* Field transfer between device and host arrays of the layouts above;
 * `count` is the number of doubles. */
int picard_field_read(picard_ctx *ctx, int32_t which, double *host, int64_t count);
int picard_field_write(picard_ctx *ctx, int32_t which, const double *host, int64_t count);

/* Kernel timing of the Jacobi solve kernel (the FP64 pass, or the split
 * form's sweep launches, two sweeps each where they apply), for roofline
 * measurement; no reference counterpart (the reference has no timers).  While
 * enabled one launch in 16 carries start/stop events stamped by its dispatch
 * packet (events around every launch would slow the timed solve; in the split
 * form the timed position moves from one refinement step to the next, so
 * launches with and without the stop test are sampled in proportion);
 * picard_timing_read returns the summed kernel time and the count of timed
 * launches since the last enable. */
int picard_timing_enable(picard_ctx *ctx, int32_t enable);
int picard_timing_read(picard_ctx *ctx, double *jacobi_ms, int64_t *jacobi_launches);
/* Jacobi sweeps the timed launches of picard_timing_read performed (a split
 * launch runs one, two or three).  No reference counterpart. */
int picard_timing_sweeps(picard_ctx *ctx, int64_t *jacobi_sweeps);
/* The same timed launches measured by the kernels themselves: the GPU's
 * 100 MHz real-time clock read by each block when it starts and when it has
 * finished, from the earliest start to the latest end (the dispatch-packet
 * events of picard_timing_read add the packet processing around the kernel).
 * No reference counterpart. */
int picard_timing_kernel_clock(picard_ctx *ctx, double *jacobi_ms, int64_t *jacobi_launches);
/* Wall time of the CalcEpotential phase of picard_step (HIP events before and
 * after the solve on the context's stream), summed over the steps run since
 * the last picard_timing_enable, and the number of solves timed.  No
 * reference counterpart (the reference calls no timer). */
int picard_timing_solve(picard_ctx *ctx, double *solve_ms, int64_t *solves);
/* FP32-split solve: `inner` FP32 sweeps per FP64 refinement step, 2..512
 * (default 512); other values return PICARD_ERR_ARG.  No reference
 * counterpart; the solution agrees with the FP64 form's to the split form's
 * tolerance for each choice.  Near the FP32 floor of a tight tolerance
 * (relative residual far below 1e-6) the solve may run up to `inner` sweeps
 * past the FP64 form's stop before the next FP64 refinement. */
int picard_split_solver(picard_ctx *ctx, int32_t inner);
/* FP32-split solve: at most this many inner sweeps per kernel launch, 1, 2
 * or 3 (default 2).  With 2, one launch performs two sweeps with the traffic
 * of one, on grids whose rows are at most 256 cells or whose x boundaries are
 * not periodic, and with an even inner length (otherwise one sweep per
 * launch); with 3, launches of three sweeps replace those of two on grids
 * whose rows are at most 256 cells with x not periodic, except the launches
 * that apply the stop test (two sweeps); on MI355X at 256^3 they take 22.6 us
 * per sweep against 18.9 with two.  The iterates are the same bit for bit.  The stop rule is then applied about every 32nd iterate while the
 * residual, extrapolated from its decay, is far from the threshold, and to
 * every second iterate near it, so the solve stops at most a sweep after the
 * first iterate that passes unless the residual falls much faster than it
 * did between tests.  No reference counterpart. */
int picard_split_sweeps_per_launch(picard_ctx *ctx, int32_t sweeps);

/* One whole main-loop body after BCinflow (picard.f90:564-631, no output
 * step): AdvancePosition, append `produced` (ParticlesProduce's particles,
 * may be 0), BCoutflow, ParticlesTransfer, ParticlesGridify, BCpotentials,
 * CalcEpotential, CalcEfield, BCfields, AdvanceVelocity. */
typedef struct {
  int64_t num_local;         /* after BCoutflow */
  int64_t removed;           /* BCoutflow removals */
  int64_t sweeps;            /* Jacobi sweeps */
  int64_t pushed;            /* particles given to AdvanceVelocity */
} picard_step_stats;

int picard_step(picard_ctx *ctx, double dt, double maxerr, int32_t maxit,
                const double *produced, const int32_t *produced_species, int64_t nproduced,
                picard_step_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
