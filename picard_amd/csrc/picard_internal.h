// picard_internal.h -- device context and kernel launchers shared by the
// translation units of libpicard_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/picard_hip.h"

namespace picard {

void set_error(const std::string &msg);

#define PICARD_HIP(call)                                                            \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::picard::set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
      return PICARD_ERR_HIP;                                                        \
    }                                                                               \
  } while (0)

#define PICARD_TRY(expr)            \
  do {                              \
    int rc_ = (expr);               \
    if (rc_ < 0) return rc_;        \
  } while (0)

// Geometry constants passed by value to every grid kernel.
struct GridDims {
  int nx, ny, nz;            // interior cells
  int X, Y, Z;               // with guards: n + 2
  int64_t G;                 // X*Y*Z nodes
  int S;                     // species
  int flow[3];
  double lmin[3], lmax[3], gmin[3], gmax[3], dxyz[3];
  double dV;
};

// Particle arrays, structure of arrays.  r[0..2] position, r[3..5] velocity,
// r[6], r[7] the reference's slots 7 and 8 (carried by every move).
struct ParticlesF64 {
  double *r[8];
  int32_t *sp;
};

// FP32-split particles: lower-node coordinates, in-cell offsets, velocities.
struct ParticlesSplit {
  int16_t *c[3];
  float *f[3];
  float *v[3];
  uint8_t *sp;
  uint32_t *id;
};

// Stop state of a Jacobi solve as seen after one launch: kernel number s reads
// slot (s-1)&1 and block 0 writes slot s&1, so no kernel reads a slot that
// one of its own blocks writes.
struct StopSlot {
  int32_t done;
  int32_t near;              // split form: the stop is expected before the next sparse test
  int64_t result;            // sweeps of the final iterate, valid when done
  double rlast;              // split form: the last tested residual norm (0: none yet)
  int64_t itlast;            // split form: the iterate it belongs to
  int32_t buf;               // split form: which correction buffer holds iterate `result`
  int32_t reserved;
};

// Device-side scalars of the Jacobi solve (one per context).
struct JacobiState {
  StopSlot slot[2];
  double bsum;               // sum of b^2 over the interior
  double thresh;             // maxerr * sqrt(bsum)
  double rsum;               // last residual sum
  int64_t pass;              // FP64 form: next pass p; split form: sweeps folded into u
  int64_t result;            // sweeps performed, valid when done
  int64_t last;              // 2*maxit + 1
  int64_t inner;             // split form: inner FP32 sweep count of the current outer step
  int32_t maxit;
  int32_t done;
  int32_t stop_inner;
  uint32_t arrive;           // block arrival counter for the last-block reduction
};

struct DeviceScratch {
  // generic
  int32_t *flags = nullptr;        // per particle
  int32_t *scan = nullptr;         // per particle (+1)
  int32_t *work[6] = {};           // per particle each: outflow rotation/stream buffers
  int32_t *scan_tmp = nullptr;     // block sums of the recursive scan
  int64_t scan_tmp_len = 0;
  int32_t *counters = nullptr;     // small device counters (errors, totals)
  // deposit (reference order)
  int32_t *key = nullptr;          // per particle
  int32_t *perm = nullptr;         // per particle
  int32_t *hist = nullptr;         // S*K + 1
  int32_t *cursor = nullptr;       // S*K
  // staging for uploads/downloads (AoS)
  double *stage = nullptr;
  int32_t *stage_sp = nullptr;
  int64_t stage_cap = 0;
  // reductions
  double *partials = nullptr;
  int64_t partials_len = 0;
};

}  // namespace picard

struct picard_ctx {
  picard_geometry geo;
  picard_species spc;
  picard::GridDims gd;
  int device = 0;
  hipStream_t stream = nullptr;

  int64_t cap = 0;           // max particles
  int64_t n = 0;             // num_local (authoritative on the host)
  picard::ParticlesF64 p{};
  bool split = false;        // PICARD_FP32_SPLIT
  // FP64 form with PICARD_DEPOSIT_FIXED_POINT: fixed-point deposit and FFT
  // face sums (order-free, within 1e-12 of the reference order)
  bool order_free = false;
  picard::ParticlesSplit ps{}, ps2{};
  unsigned long long *Nfix = nullptr;   // split: fixed-point densities
  float *E32 = nullptr;      // split: E as float4 per node
  // split solve, compact interior layout (rows padded to 4 cells): right-hand
  // side, ping-pong correction, inner iterates j % 3 on periodic boundary planes
  float *r32 = nullptr, *e32[2] = {nullptr, nullptr};
  float *eo32[4] = {nullptr, nullptr, nullptr, nullptr};
  int mx_inner = 512;        // split: FP32 sweeps per FP64 refinement
  bool mx_pairs = true;      // split: two sweeps per launch where the grid allows
  bool mx_triples = false;   // split: three sweeps per launch where the grid allows
  int pair_wmax = 0;         // split: resident blocks per XCD of the two-sweep kernel
  // split: the two-sweep kernel that shares rows through LDS where the grid
  // allows (x not periodic, nx <= 256); PICARD_PAIR_KERNEL=regs at create
  // selects the register-window kernel everywhere
  bool pair_lds = true;
  int lds_wmax = 0;          // split: resident blocks per XCD of the LDS two-sweep kernel
  int tri_wmax = 0;          // split: resident blocks per XCD of the three-sweep kernel
  int outer_wmax = 0;        // split: resident blocks per XCD of the outer-step kernel
  double *Uw = nullptr;      // split: second FP64 potential buffer of the solve
  int sweep_wmax = 0;        // split: resident blocks per XCD of the sweep kernel (0: not yet known)
  int pass_wmax = 0;         // FP64: the same for the Jacobi pass kernel
  int pass2_wmax = 0;        // FP64: the same for the two-cells-per-lane pass kernel
  bool bcpot_lds_set = false;  // FP64: the LDS face-sum kernel's LDS limit has been raised
  int64_t next_id = 0;

  // grids
  double *U = nullptr, *E = nullptr, *N = nullptr, *Nraw = nullptr, *C = nullptr, *D = nullptr;
  double *jb[3] = {nullptr, nullptr, nullptr};   // Jacobi iterate buffers (G nodes each)
  double *b = nullptr;
  double *jb_mem[3] = {nullptr, nullptr, nullptr};   // their allocations: jb[q] = jb_mem[q] + 1
  double *b_mem = nullptr;                           // b = b_mem + 1
  double *consts = nullptr;  // device: n2p[8], charge[8]
  int64_t K = 0;             // deposit cells per species: (nx+1)(ny+1)(nz+1)
  bool have_D = false;

  // split form: FFT evaluation of the open-boundary face sums
  double2 *fft_K[3][2] = {};     // per flow dim and face: transformed face kernels, nd planes
  double2 *fft_tw[3][2] = {};    // per flow dim: twiddles of the two transverse lengths
  double2 *fft_work = nullptr;   // transformed charge planes + two face accumulators
  int64_t fft_work_len = 0;
  bool fft_ready = false;

  // number fluxes of ParticlesGridifyAll (allocated on first use)
  double *F = nullptr;                 // F(species, 3, node)
  double *Fraw = nullptr;              // FP64 form: before the guard folds
  unsigned long long *Ffix = nullptr;  // split form: fixed-point accumulators

  // device Haser production (ParticlesProduce) run by picard_step
  int32_t *haser_off = nullptr;  // per (cell, producing species): slot counts, then offsets
  int64_t haser_len = 0;
  bool prod_on = false;
  picard_plasma prod_plasma{};
  picard_haser prod_haser{};
  uint64_t prod_seed = 0;
  int64_t steps_done = 0;        // picard_step calls completed

  picard::JacobiState *js = nullptr;   // device
  picard::JacobiState *js_host = nullptr;   // pinned mirror
  int32_t *host_counters = nullptr;    // pinned mirror of scratch.counters

  picard::DeviceScratch s;

  // solve-to-solve predictor of the Jacobi sweep count (sizes the first batch)
  int64_t last_sweeps = 0;

  // kernel timing of the Jacobi sweep (picard_timing_enable): event pairs
  // stamped by the dispatch packets themselves, around one sweep in
  // picard::kTimeEvery
  bool timing = false;
  std::vector<hipEvent_t> tev;
  // kernel-duration stamps of the sampled solve launches (2 x 16 words per
  // launch, in ticks of the 100 MHz real-time clock) and their host copy
  static constexpr int kStamps = 8192;
  static constexpr int kStampWords = 32;
  unsigned long long *kst = nullptr;
  std::vector<unsigned long long> kst_host;
  double t_ms = 0.0;
  int64_t t_launches = 0;
  int64_t t_sweeps = 0;      // Jacobi sweeps of the timed launches
  double tk_ms = 0.0;        // the same launches by the kernels' own clock stamps
  int64_t tk_launches = 0;
  // wall time of the whole CalcEpotential phase inside picard_step while
  // timing is on (events before and after the solve on the context's stream)
  hipEvent_t solve_ev[2] = {nullptr, nullptr};
  double solve_ms = 0.0;
  int64_t solves = 0;
};

namespace picard {

constexpr int kTimeEvery = 16;   // timed sweeps: one launch in this many
// Longest FP32-split refinement step (FP32 sweeps per FP64 refinement) the
// API accepts: with one refinement covering a whole small solve (4096 sweeps
// at 16^3, a 1,832-sweep solve) the FP32 iterate stopped 26 sweeps after the
// FP64 form, outside the agreement the split form keeps for 2..512.
constexpr int kMaxSplitInner = 512;

// ---- particles.hip ----
int launch_push_x(picard_ctx *c, int64_t first, int64_t count, double dt);
int launch_outflow(picard_ctx *c, int64_t *n_new, int64_t *removed);
int launch_wrap(picard_ctx *c);
// mode 0: AdvanceVelocity; 1: AdvanceVelocityImp; 2: AdvanceVelocityExp (deltat)
int launch_boris(picard_ctx *c, double dt, int mode = 0, double deltat = 0.0);
int launch_append(picard_ctx *c, const double *host_rp, const int32_t *host_sp, int64_t count);
int launch_read(picard_ctx *c, int64_t first, int64_t count, double *host_rp, int32_t *host_sp);
// Reset the first n kernel-duration stamp pairs; read back n and add their
// durations to the timing totals.
int kstamp_reset(picard_ctx *c, int n);
int kstamp_collect(picard_ctx *c, int n);
// particles [first, first + count) as real_particles(8, n) / int_particles(n)
// into device buffers (either form)
int stage_particles(picard_ctx *c, int64_t first, int64_t count, double *stage, int32_t *stage_sp);

int ensure_stage(picard_ctx *c, int64_t want);

// ---- grid.hip ----
int launch_bnorm(picard_ctx *c, double maxerr);
// flux: also the number fluxes F of ParticlesGridifyAll
int launch_gridify(picard_ctx *c, bool flux = false);
int launch_calc_distance(picard_ctx *c);
int launch_bc_potentials(picard_ctx *c);
int launch_calc_epotential(picard_ctx *c, double maxerr, int32_t maxit, int64_t *sweeps);
int launch_calc_efield(picard_ctx *c);
int launch_bc_fields(picard_ctx *c);

// ---- diag.hip ----
int diagnostics(picard_ctx *c, const double *mass, picard_diag *out);

// ---- bcpot_fft.hip ----
int bcpot_fft_prepare(picard_ctx *c);
int launch_bc_potentials_fft(picard_ctx *c);
void bcpot_fft_release(picard_ctx *c);

// ---- sources.hip ----
int generate_initial(picard_ctx *c, const picard_plasma *pl, uint64_t seed);
int generate_inflow(picard_ctx *c, const picard_plasma *pl, uint64_t seed, int64_t step);
int produce_haser(picard_ctx *c, const picard_plasma *pl, const picard_haser *hs, uint64_t seed,
                  int64_t step, double dt, int64_t *produced);

// ---- split.hip ----
int sp_push(picard_ctx *c, double dt);
int sp_sort(picard_ctx *c, int wrap);
int sp_gridify(picard_ctx *c, bool flux = false);
// allocates the flux arrays (F, and Fraw or Ffix) on first use
int ensure_flux(picard_ctx *c);
int sp_epotential(picard_ctx *c, double maxerr, int32_t maxit, int64_t *sweeps);
int sp_efield(picard_ctx *c);
int sp_widen_e(picard_ctx *c);
int sp_boris(picard_ctx *c, double dt, int mode = 0, double deltat = 0.0);
int sp_append(picard_ctx *c, const double *host_rp, const int32_t *host_sp, int64_t count);
int sp_read(picard_ctx *c, int64_t first, int64_t count, double *host_rp, int32_t *host_sp);
int sp_stage(picard_ctx *c, int64_t first, int64_t count, double *stage, int32_t *stage_sp);

// ---- scan.hip ----
// Exclusive prefix sum of n int32 values (in may alias out); out[n] = total.
int scan_exclusive_i32(picard_ctx *c, const int32_t *in, int32_t *out, int64_t n);
// The same on a stream with a caller's temporary arena (n >= 1, captured into
// graphs: no memset).
int scan_exclusive_stream(hipStream_t st, const int32_t *in, int32_t *out, int64_t n,
                          int32_t *tmp, int64_t tmp_len);
int64_t scan_tmp_required(int64_t n);

// 3D launch shape over the interior cells: x fastest, 256 threads per block.
struct Launch3 {
  dim3 grid, block;
  int64_t blocks() const { return (int64_t)grid.x * grid.y * grid.z; }
};

inline Launch3 interior_launch(int nx, int ny, int nz) {
  int bx = 64;
  while (bx > 8 && bx / 2 >= nx) bx /= 2;
  const int by = 256 / bx;
  Launch3 L;
  L.block = dim3(bx, by, 1);
  L.grid = dim3((nx + bx - 1) / bx, (ny + by - 1) / by, nz);
  return L;
}

inline int grid_blocks(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace picard
