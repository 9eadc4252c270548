/* picard_ensemble.h -- many independent Picard simulations advanced together
 * on one MI355X, in the FP32-split form.
 *
 * No reference counterpart: the reference runs one simulation per process
 * (picard.f90:12-684).  Each simulation here is one pass of that main loop per
 * step (picard.f90:554-633: BCinflow, AdvancePosition, BCoutflow,
 * ParticlesTransfer, ParticlesGridify, BCpotentials, CalcEpotential,
 * CalcEfield, BCfields, AdvanceVelocity), with the boundaries of the 32^3
 * workload: x in/outflow, y and z periodic, H+ and e- at `ppc` macroparticles
 * per cell each.  Every phase of a step is one kernel launch over a work list
 * of (simulation, tile) items covering every simulation, and the phases are
 * replayed from hipGraphs captured once.  Simulations join (picard_ensemble_add)
 * and leave (picard_ensemble_remove) between steps; their particles and fields
 * stay in HBM for their whole life.
 *
 * Per simulation, a step's results do not depend on which other simulations
 * share the launches, on the order of the slots or on the tile shapes the
 * autotuner picks: deposits are int64 fixed point, the stop-test residual norm
 * of the Jacobi solve is accumulated in fixed point, and the open-boundary face
 * sums reduce their partials in a fixed order.
 *
 * The FP32-split form: particles as int16 cell + FP32 offset with FP32
 * velocities, FP32 Jacobi sweeps on the correction with FP64 refinement every
 * `inner` sweeps, FP64 face sums (BCpotentials) and FP64 refinement residuals.
 * The stop rule is CalcEpotential.f90:60-62 (relative residual maxerr, tested
 * before every update, at most 2*maxit+1 sweeps) applied to every iterate, with
 * fresh periodic images in the residual. */
#ifndef PICARD_ENSEMBLE_H
#define PICARD_ENSEMBLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One simulation's parameters (the run configuration of picard.f90:140-302
 * restricted to the ensemble's plasma): a box of nx*ny*nz cubic cells of size
 * dx centred on the origin, H+ of number density `density` with as many
 * electrons, both at `kelvin`, drifting at `drift` along +x, uniform B,
 * dt = dt_wpe / omega_pe, the Jacobi tolerance and iteration cap, and the key
 * of the simulation's counter-based random streams. */
typedef struct {
  int32_t nx, ny, nz;
  int32_t ppc;              /* H+ macroparticles per cell (each with one e-) */
  double dx;                /* cell size [m] */
  double density;           /* H+ number density [m^-3] */
  double kelvin;            /* temperature of both species [K] */
  double drift;             /* bulk velocity along +x [m/s] */
  double B[3];              /* [T] */
  double dt_wpe;            /* time step in units of 1/omega_pe */
  double maxerr;            /* Jacobi relative tolerance */
  int32_t maxit;            /* Jacobi iteration cap (2*maxit+1 sweeps at most) */
  int32_t reserved;
  uint64_t seed;            /* key of the simulation's Philox streams */
} picard_sim_params;

/* The constants picard.f90:333-466 derives for such a simulation (species
 * 1 = e-, 2 = H+): cell size and volume, macroparticle weights, charge over
 * mass, the mass-weighted bulk velocity, dt, omega_pe, and the Boris tables
 * tB(1:3, s), sB(1:3, s) with |tB| in tB4. */
typedef struct {
  double dxyz, dV;
  double lmin[3], lmax[3];
  double charge[2], n2p[2], eta[2];
  double ve[3];
  double dt, omega_pe;
  double tB[2][3], tB4[2], sB[2][3];
} picard_sim_derived;

/* Host only (no device work): the derived constants of one simulation.
 * Returns PICARD_ERR_ARG for sizes or values outside what the ensemble runs. */
int picard_sim_derive(const picard_sim_params *p, picard_sim_derived *out);

/* Device memory a simulation holds while it is in an ensemble (grids, its
 * share of the sort keys and the face-sum scratch), and the particle slots it
 * needs at its initial load plus one step's inflow.  Host only. */
int picard_sim_footprint(const picard_sim_params *p, int64_t *grid_bytes, int64_t *particles);

typedef struct {
  int32_t device;           /* HIP device ordinal */
  int32_t max_sims;         /* slots */
  int64_t max_cells;        /* interior cells over all simulations at once */
  int64_t max_particles;    /* particles over all simulations at once */
  int32_t inner;            /* FP32 sweeps per FP64 refinement: 64..512, a multiple of 64 (0: 512) */
  int32_t use_graphs;       /* nonzero: replay the step's phases from captured hipGraphs */
} picard_ensemble_config;

typedef struct picard_ensemble picard_ensemble;

/* Tile shape of the solve kernels: cells per thread (1, 2, 4 or 8), threads
 * per block (256), resident blocks per CU of the persistent launch (1..8).  A
 * tile is threads * cells consecutive interior cells of one simulation, thread
 * t taking cells t, t + threads, ...; the autotuner tries the twelve
 * combinations of cells per thread and blocks per CU in {2, 4, 8}. */
typedef struct {
  int32_t cells_per_thread;
  int32_t threads;
  int32_t blocks_per_cu;
  int32_t reserved;
} picard_ensemble_tile;

typedef struct {
  int32_t sims;             /* simulations stepped */
  int32_t noconv;           /* of them, solves that ran 2*maxit+1 sweeps */
  int64_t particles;        /* particles after the step, all simulations */
  int64_t pushed;           /* particles given to AdvanceVelocity, all simulations */
  int64_t max_sweeps;       /* largest Jacobi sweep count of the step */
  int64_t sum_sweeps;       /* sum of the simulations' sweep counts */
  int64_t solve_chunks;     /* solve graph replays (64 sweeps each) */
  int64_t cell_sweeps;      /* sum over simulations of interior cells x sweeps */
  double step_ms;           /* host wall time of the step */
  double solve_ms;          /* of it, the solve (refinement passes and sweeps) */
} picard_ensemble_stats;

/* Per-simulation status after the last step. */
typedef struct {
  int32_t live;             /* the slot holds a simulation */
  int32_t noconv;           /* its last solve ran 2*maxit+1 sweeps */
  int64_t steps;            /* steps completed */
  int64_t num_local;        /* particles after the last step */
  int64_t pushed;           /* particles pushed at the last step */
  int64_t sweeps;           /* Jacobi sweeps of the last step */
} picard_sim_status;

int picard_ensemble_create(const picard_ensemble_config *cfg, picard_ensemble **out);
int picard_ensemble_destroy(picard_ensemble *e);

/* Admit a simulation into a free slot: reserves its grids and keys, builds
 * its distance table (CalcDistance.f90:25-91) and loads its initial plasma
 * (ParticlesInitialize.f90:54-119 on its Philox streams).  Returns
 * PICARD_ERR_CAPACITY when no slot, cells or particle room is left. */
int picard_ensemble_add(picard_ensemble *e, const picard_sim_params *p, int32_t *slot);
/* Retire a simulation: its particles are dropped by the next step's sort and
 * its grids and keys are freed. */
int picard_ensemble_remove(picard_ensemble *e, int32_t slot);

/* One step of every live simulation. */
int picard_ensemble_step(picard_ensemble *e, picard_ensemble_stats *stats);

int picard_ensemble_status(picard_ensemble *e, int32_t slot, picard_sim_status *out);
/* A simulation's fields in the reference layout (see picard_hip.h): which =
 * PICARD_FIELD_U, PICARD_FIELD_E (widened from FP32), PICARD_FIELD_N or
 * PICARD_FIELD_C; count is the element count. */
int picard_ensemble_field_read(picard_ensemble *e, int32_t slot, int32_t which, double *host,
                               int64_t count);
/* A simulation's particles after the last step as real_particles(8, n) /
 * int_particles(n), slot 8 holding each particle's id; count must equal its
 * num_local. */
int picard_ensemble_particles_read(picard_ensemble *e, int32_t slot, double *rp, int32_t *sp,
                                   int64_t count);

/* Autotuner: times `sweeps` solve sweeps of the current simulations under
 * every candidate tile shape (results are discarded: the fields are restored)
 * and keeps the fastest; `times_us` (may be NULL) receives up to `max_times`
 * per-sweep times in candidate order, `chosen` the shape kept. */
int picard_ensemble_tune(picard_ensemble *e, int32_t sweeps, picard_ensemble_tile *chosen,
                         double *times_us, int32_t max_times, int32_t *ncandidates);
int picard_ensemble_set_tile(picard_ensemble *e, const picard_ensemble_tile *tile);
int picard_ensemble_get_tile(picard_ensemble *e, picard_ensemble_tile *tile);

#ifdef __cplusplus
}
#endif
#endif
