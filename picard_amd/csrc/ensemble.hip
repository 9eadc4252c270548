// ensemble.hip -- many independent simulations stepped together in the
// FP32-split form (include/picard_ensemble.h).
//
// Layout.  All simulations share one particle arena (structure of arrays,
// double-buffered for the sort) and one arena per grid quantity; a slot owns
// a region of each grid arena (first-fit, freed when it leaves) and a range
// of sort keys.  Every step the particles of all simulations are sorted
// together by (slot key range, lower node, species), so each simulation's
// particles are contiguous after the sort, in the order of its key range.
//
// Work lists.  Every phase is one launch of a fixed grid (persistent, but for
// the scatter's, sized for the particle arena's capacity) that walks a
// device-resident list of (slot, tile) items: node tiles for the per-node
// kernels, cell tiles (the autotuned shape) for the solve, face tiles for the
// open-boundary sums, and the particle range for the particle kernels.  The lists and their lengths live in device memory, so the
// captured graphs stay valid when simulations join or leave.
//
// Solve.  Mixed precision per simulation: an FP64 refinement pass forms
// u' = u + e and r = b - A u' (ping-pong FP64 buffers), then FP32 Jacobi
// sweeps on A e = r.  Every sweep launch also evaluates the residual of the
// iterate it reads, r - A e_j, and converts each cell's square to fixed point
// (units of the simulation's threshold^2 * 2^-40, clamped at 2^40 + 1); each
// tile stores its integer sum, and a small launch after every kDecideEvery
// sweeps sums each simulation's tiles (integers: the same for any tiling and
// any mix of simulations) for all of those iterates in one pass and applies
// the stop rule of CalcEpotential.f90:60-62 (sum <= 2^40, i.e. |res| <=
// maxerr |b|) to each in turn, stopping at the first that passes, advancing
// or stopping that simulation.  The
// corrections rotate over kDecideEvery + 1 buffers, so the iterate the rule
// picks is still held when it is picked: a simulation stops on the same
// iterate as with a decision after every sweep, and the sweeps it ran past it
// are discarded.  Simulations stop independently; tiles of stopped ones are
// skipped.  (One atomic per tile on a per-simulation sum and counter cost
// 0.76 ms per sweep for 64 simulations: the tiles of one simulation met on the
// same address.  A decision launch after every sweep cost 4 us of kernel and a
// launch gap per sweep.)
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/picard_ensemble.h"
#include "picard_internal.h"
#include "reduce.h"
#include "rng.h"

namespace picard {
namespace ens {

constexpr int kT = 256;
constexpr double kFix = 1099511627776.0;      // 2^40: deposit weights
constexpr double kFixInv = 1.0 / 1099511627776.0;
constexpr unsigned long long kStopUnits = 1ull << 40;   // the threshold in residual units
constexpr double kTau = 6.2831853071795864769252867665590;
constexpr double kEC = 1.60217646e-19, kMP = 1.6726216e-27, kME = 9.1093819e-31;
constexpr double kEPS0 = 8.854187817e-12;
constexpr int kChunkSweeps = 64;     // sweeps per replay of the solve graph
constexpr int kDecideEvery = 8;      // sweeps per stop decision (divides kChunkSweeps)
constexpr int kEBufs = kDecideEvery + 1;   // correction buffers
constexpr int kFacePlanes = 8;       // x-planes per face-sum tile
constexpr int kMaxSlots = 4096;

// ---------------------------------------------------------------------------
// device tables

struct SimDev {
  int32_t nx, ny, nz, live;
  int32_t X, Y, Z, ppc;
  int64_t G, gbase;         // guard-layout nodes and their offset in U, u0/u1, C, E32 (2x in N, Nfix)
  int64_t cells, cbase;     // interior cells and their offset in r32, e32
  int64_t dbase;            // distance table D(a, b, c), b fastest: (nx+1)(ny+1)(nz+1)
  int64_t fbase;            // face-sum partials: 2 faces x fchunks x ny*nz
  int32_t kbase, nkeys;     // sort keys
  int32_t tiles, fchunks;   // solve tiles (current shape), x-plane chunks of the face sums
  int64_t ubase;            // first entry of this slot's per-tile residual units
  int64_t pfirst;           // first particle of this slot in one step's inflow (live slots)
  int64_t inflow;           // particles one step's inflow appends for this slot
  double lmin[3], xmax;     // box lower corner, Lx_max
  double dx, dVeps0, kappa; // cell size, dV * eps0, 1 / (4 pi eps0)
  double n2p[2], charge[2];
  double i2;                // 1 / dx^2
  double maxerr;
  int32_t maxit, pad0;
  float etadt2[2], tB[2][3], sB[2][3], ve[3], dtdx;
  // FP32 sweep: e' = wf sum6 + kf e - aff r with 6 wf + kf = 1 exactly (wf =
  // fl32(1/6), kf = 1 - 6 wf = -2^-25), so the iteration keeps the FP64
  // Jacobi spectrum: with 6 fl32(1/6) alone both ends move by ~1e-7, which
  // near the slowest modes' decay changes the stop by many sweeps;
  // res = r - i2f (sum6 - 6 e)
  float wf, kf, aff, i2f;
  uint64_t seed;
  PlasmaConst pc;
};

struct SimState {
  int32_t j, solving, noconv, have_e;
  int32_t ucur, stop_j, pad0, pad1;
  int64_t base, sweeps, step, next_id;
  int64_t start, count, pushed;
  double thresh2, ifixq;    // threshold^2 and 2^40 / threshold^2 (residual units per V^2/m^4)
};

struct Global {
  int64_t n;                // particles in the current arena
  int64_t inflow_total;     // particles one step's inflow appends
  int32_t n_node_tiles, n_cell_tiles, n_face_tiles, n_fred_tiles;
  int32_t n_live, n_inflow_map;
  int32_t active;           // simulations still solving
  int32_t err;              // bit 0: particle capacity; bit 1: particle outside its grid
  int32_t chunk_j0;         // inner sweep index of the solve chunk's first sweep
  int32_t pad;
};

struct Tile {               // work-list item
  int32_t slot, a, b, c;
};

struct PartArena {
  int16_t *c[3];
  float *f[3];
  float *v[3];
  uint8_t *sp;
  uint16_t *sim;
  uint32_t *id;
};

struct Arenas {
  unsigned long long *units;   // per solve tile and iterate of the decision group:
                               // fixed-point residual units, [j % kDecideEvery][tile]
  int64_t ucap;             // tiles per units slot
  SimDev *sims;
  SimState *state;
  Global *g;
  Tile *node_tiles, *cell_tiles, *face_tiles, *fred_tiles;
  int32_t *live;            // live slots
  Tile *inflow_map;         // {slot, first ion, 0, 0} for live slots in ion order
  double *U, *u[2], *C, *N, *D, *fscr;
  unsigned long long *Nfix;
  float4 *E32;
  float *r32, *e32[kEBufs];
  int32_t *hist, *cursor, *key;
  int64_t kcap;             // key capacity
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// ---------------------------------------------------------------------------
// sources: ParticlesInitialize.f90:54-119 (interior) and BCinflow.f90:53-191
// (the two x faces; y and z are periodic) on the slot's Philox streams, with
// the single-domain form's keys, tags, slot order and ids.

__device__ __forceinline__ void put_split(const SimDev &sd, PartArena p, int64_t q, uint32_t id,
                                          int sp, int slot, double x, double y, double z,
                                          const double v[3]) {
  const double r[3] = {x, y, z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xi = (r[d] - sd.lmin[d]) / sd.dx + 0.5;
    double fl = floor(xi);
    float f = (float)(xi - fl);
    if (f >= 1.0f) {
      f = 0.0f;
      fl += 1.0;
    }
    p.c[d][q] = (int16_t)fl;
    p.f[d][q] = f;
    p.v[d][q] = (float)v[d];
  }
  p.sp[q] = (uint8_t)sp;
  p.sim[q] = (uint16_t)slot;
  p.id[q] = id;
}

// One H+ ion (and its electron) of slot `slot`: cell (ci, cj, ck) 1-based,
// lower edge x0 given, stream counter `local` within its block, tag.
__device__ __forceinline__ void emit_ion(const SimDev &sd, PartArena p, int slot, int64_t q,
                                         uint32_t id, uint64_t local, uint32_t tag, double x0,
                                         double y0, double z0) {
  Stream st{sd.seed, (uint32_t)local, (uint32_t)(local >> 32), tag};
  const double px = x0 + sd.dx * st.uniform();
  const double py = y0 + sd.dx * st.uniform();
  const double pz = z0 + sd.dx * st.uniform();
  double v[3];
  maxwellian(st, sd.pc, 1, 1, v);
  half_rotate(sd.pc, 1, v);
  put_split(sd, p, q, id, 2, slot, px, py, pz, v);
  double ve_[3];
  maxwellian(st, sd.pc, 0, 1, ve_);
  half_rotate(sd.pc, 0, ve_);
  put_split(sd, p, q + 1, id + 1, 1, slot, px, py, pz, ve_);
}

// Initial load of one slot at arena position `at`: interior cells x fastest,
// ppc ions per cell, each followed by its electron (tag 0xA11 as the single
// domain's generate_initial).
__global__ void __launch_bounds__(kT) k_initial(Arenas A, PartArena p, int slot, int64_t at) {
  const SimDev &sd = A.sims[slot];
  const int64_t ions = sd.cells * sd.ppc;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < ions;
       t += (int64_t)gridDim.x * kT) {
    const int64_t cell = t / sd.ppc;
    const int ion = (int)(t % sd.ppc);
    const int ci = 2 + (int)(cell % sd.nx), cj = 2 + (int)((cell / sd.nx) % sd.ny);
    const int ck = 2 + (int)(cell / ((int64_t)sd.nx * sd.ny));
    const int64_t local = cell * 2 * sd.ppc + 2 * ion;
    emit_ion(sd, p, slot, at + local, (uint32_t)local, (uint64_t)local,
             0xA11u, sd.lmin[0] + sd.dx * (ci - 2), sd.lmin[1] + sd.dx * (cj - 2),
             sd.lmin[2] + sd.dx * (ck - 2));
  }
}

// One step's inflow of every live slot, appended after the arena's particles:
// +x face (x0 = Lx_max, tag base+0) then -x face (x0 = Lx_min - dx, base+1),
// cells y fastest, tag base = 0x1000 + 8 * step (the single domain's
// generate_inflow for step = steps done + 1).
__global__ void __launch_bounds__(kT) k_inflow(Arenas A, PartArena p, int64_t cap) {
  const Global gl = *A.g;
  if (gl.n + gl.inflow_total > cap) return;          // k_inflow_commit flags it
  const int64_t ions = gl.inflow_total / 2;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < ions;
       t += (int64_t)gridDim.x * kT) {
    // the live slot whose ions hold t: last map entry with first <= t
    int lo = 0, hi = gl.n_inflow_map - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((int64_t)A.inflow_map[mid].a <= t) lo = mid;
      else hi = mid - 1;
    }
    const int slot = A.inflow_map[lo].slot;
    const SimDev &sd = A.sims[slot];
    const int64_t li = t - A.inflow_map[lo].a;
    const int64_t face_ions = (int64_t)sd.ny * sd.nz * sd.ppc;
    const int face = (int)(li / face_ions);
    const int64_t lf = li % face_ions;
    const int64_t cell = lf / sd.ppc;
    const int ion = (int)(lf % sd.ppc);
    const int cj = 2 + (int)(cell % sd.ny), ck = 2 + (int)(cell / sd.ny);
    const int64_t local = cell * 2 * sd.ppc + 2 * ion;       // slot within the face's block
    const int64_t q = gl.n + sd.pfirst + face * 2 * face_ions + local;
    const SimState &st = A.state[slot];
    const uint32_t tag = 0x1000u + 8u * (uint32_t)(st.step + 1) + (uint32_t)face;
    const double x0 = face == 0 ? sd.xmax : sd.lmin[0] - sd.dx;
    emit_ion(sd, p, slot, q, (uint32_t)(st.next_id + face * 2 * face_ions + local),
             (uint64_t)local, tag, x0, sd.lmin[1] + sd.dx * (cj - 2),
             sd.lmin[2] + sd.dx * (ck - 2));
  }
}

__global__ void k_inflow_commit(Arenas A, int64_t cap) {
  Global *g = A.g;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (g->n + g->inflow_total > cap) {
    g->err |= 1;
    return;
  }
  for (int q = 0; q < g->n_live; ++q) {
    const int s = A.live[q];
    A.state[s].next_id += A.sims[s].inflow;
  }
  g->n += g->inflow_total;
}

// AdvancePosition on the split coordinates: xi += v dt/dx.
__device__ __forceinline__ int carry(float &f) {
  const float fl = floorf(f);
  f = f - fl;
  int c = (int)fl;
  if (f >= 1.0f) {
    f = 0.0f;
    c += 1;
  }
  return c;
}

__global__ void __launch_bounds__(kT) k_push(Arenas A, PartArena p) {
  const int64_t n = A.g->n;
  for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < n; q += (int64_t)gridDim.x * kT) {
    const float dtdx = A.sims[p.sim[q]].dtdx;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float f = p.f[d][q] + p.v[d][q] * dtdx;
      p.c[d][q] = (int16_t)(p.c[d][q] + carry(f));
      p.f[d][q] = f;
    }
  }
}

// ---------------------------------------------------------------------------
// sort: BCoutflow (x faces, and the band one domain width out in every
// dimension, BCoutflow.f90:28-66) and ParticlesTransfer's periodic wrap of y
// and z (ParticlesTransfer.f90:494-519) on the split coordinates, then a
// counting sort over every slot's key range.  Particles of slots that left
// are dropped.

__device__ __forceinline__ bool below(int c, float f) { return c < 0 || (c == 0 && f < 0.5f); }
__device__ __forceinline__ bool above(int c, float f, int n) { return c > n || (c == n && f > 0.5f); }

// One atomic per distinct key and wave: the lanes sharing the leader's key
// take ranks by lane order after the value the leader's add returned.  Every
// lane stays in the loop (its condition is the wave-uniform pending mask).
__device__ __forceinline__ int32_t wave_add1(int32_t *counter, int32_t k) {
  const int lane = lane_id();
  bool pending = k >= 0;
  int32_t ret = -1;
  uint64_t todo = __ballot(pending);
  while (todo) {
    const int leader = (int)__builtin_ctzll(todo);
    const int32_t kl = __builtin_amdgcn_readlane(k, leader);
    const bool mine = pending && k == kl;
    const uint64_t m = __ballot(mine) | (1ull << leader);
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(&counter[kl], (int32_t)__popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    if (mine) {
      ret = base + (int32_t)__popcll(m & ((1ull << lane) - 1ull));
      pending = false;
    }
    todo &= ~m;
  }
  return ret;
}

__global__ void __launch_bounds__(kT) k_flag(Arenas A, PartArena p) {
  const int64_t n = A.g->n;
  int nerr = 0;
  // block-uniform trip count: every lane of a wave takes part in wave_add1
  for (int64_t b0 = (int64_t)blockIdx.x * kT; b0 < n; b0 += (int64_t)gridDim.x * kT) {
    const int64_t q = b0 + threadIdx.x;
    int32_t k = -1;
    if (q < n) {
      const int slot = p.sim[q];
      const SimDev &sd = A.sims[slot];
      if (sd.live) {
        const int nn[3] = {sd.nx, sd.ny, sd.nz};
        int c[3] = {p.c[0][q], p.c[1][q], p.c[2][q]};
        const float f[3] = {p.f[0][q], p.f[1][q], p.f[2][q]};
        bool del = below(c[0], f[0]) || above(c[0], f[0], nn[0]);      // x in/outflow
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (below(c[d] + nn[d], f[d]) || above(c[d] - nn[d], f[d], nn[d])) del = true;
        if (!del) {
#pragma unroll
          for (int d = 2; d >= 1; --d) {           // z then y (x is not periodic)
            if (below(c[d], f[d])) c[d] += nn[d];
            if (above(c[d], f[d], nn[d])) c[d] -= nn[d];
          }
          if (c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] <= nn[0] && c[1] <= nn[1] &&
              c[2] <= nn[2]) {
            p.c[1][q] = (int16_t)c[1];
            p.c[2][q] = (int16_t)c[2];
            const int32_t lin = c[0] + (sd.nx + 1) * (c[1] + (sd.ny + 1) * c[2]);
            k = sd.kbase + 2 * lin + ((int32_t)p.sp[q] - 1);
          } else {
            ++nerr;
          }
        }
      }
      A.key[q] = k;
    }
    wave_add1(A.hist, k);
  }
  if (nerr) atomicOr(&A.g->err, 2);
}

// One particle per thread over a grid sized for the arena's capacity (the
// count is in device memory, so the captured launch cannot size itself to
// it): blocks past the count leave at once, and every wave in range walks a
// single key-cursor-copy chain instead of hundreds in a grid-stride loop.
__global__ void __launch_bounds__(kT) k_scatter(Arenas A, PartArena src, PartArena dst) {
  const int64_t n = A.g->n;
  const int64_t b0 = (int64_t)blockIdx.x * kT;
  if (b0 >= n) return;                                   // block-uniform
  const int64_t q = b0 + threadIdx.x;
  const int32_t k = q < n ? A.key[q] : -1;               // lanes past n take part in the adds
  const int32_t r = wave_add1(A.cursor, k);
  if (k < 0 || r < 0) return;
  const int64_t at = (int64_t)A.hist[k] + r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dst.c[d][at] = src.c[d][q];
    dst.f[d][at] = src.f[d][q];
    dst.v[d][at] = src.v[d][q];
  }
  dst.sp[at] = src.sp[q];
  dst.sim[at] = src.sim[q];
  dst.id[at] = src.id[q];
}

// Each live slot's particle range after the scatter; the arena's new count.
__global__ void k_ranges(Arenas A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const Global &gl = *A.g;
  if (q < gl.n_live) {
    const int s = A.live[q];
    const SimDev &sd = A.sims[s];
    const int64_t a = A.hist[sd.kbase], b = A.hist[sd.kbase + sd.nkeys];
    A.state[s].start = a;
    A.state[s].count = b - a;
  }
  if (q == 0) A.g->n = A.hist[A.kcap];
}

// ---------------------------------------------------------------------------
// deposit (ParticlesGridify.f90:59-121): int64 fixed point, runs of one
// (slot, cell, species) summed in registers, one atomic per corner and run.

constexpr int kDepChunk = 16;
constexpr int kDepSeg = 128;          // nodes per LDS segment of the deposit

// A thread adds the fixed-point weights of each run of one (cell, species)
// among its kDepChunk sorted particles in registers and flushes one add per
// corner when the run ends.  The adds are staged in LDS per block chunk of
// kT * kDepChunk particles: from the chunk's first particle (slot s0, lower
// node b0) the corners of runs of the same slot with lower nodes in [b0, b0 +
// kDepSeg - 1) fall in four node segments b0 + {0, X, XY, XY + X} + [0,
// kDepSeg), added with LDS atomics; other runs go straight to memory.  The
// block then adds each nonzero segment entry to memory once.  Integer sums:
// the same result as one atomic per particle.
__global__ void __launch_bounds__(kT) k_deposit(Arenas A, PartArena p) {
  __shared__ unsigned long long seg[4 * kDepSeg * 2];
  __shared__ int64_t seg_b0;
  __shared__ int seg_slot;
  const int64_t n = A.g->n;
  const int64_t span = (int64_t)kT * kDepChunk;
  for (int64_t cb = (int64_t)blockIdx.x * span; cb < n; cb += (int64_t)gridDim.x * span) {
    if (threadIdx.x == 0) {
      int slot0 = -1;
      int64_t bb = 0;
      const int sl = p.sim[cb];
      const SimDev &sd = A.sims[sl];
      const int i = p.c[0][cb], j = p.c[1][cb], k = p.c[2][cb];
      if (!(i < 0 || j < 0 || k < 0 || i > sd.nx || j > sd.ny || k > sd.nz)) {
        slot0 = sl;
        bb = sd.gbase + i + (int64_t)sd.X * j + (int64_t)sd.X * sd.Y * k;
      }
      seg_slot = slot0;
      seg_b0 = bb;
    }
    for (int e = threadIdx.x; e < 4 * kDepSeg * 2; e += kT) seg[e] = 0;
    __syncthreads();
    const int slot0 = seg_slot;
    const int64_t b0 = seg_b0;
    const int64_t q0 = cb + (int64_t)threadIdx.x * kDepChunk;
    const int64_t q1 = q0 + kDepChunk < n ? q0 + kDepChunk : n;    // empty past n
    int64_t cur = -1;
    int cur_s = 0, cur_slot = -1;
    int64_t X = 0, XY = 0;
    unsigned long long W[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto flush = [&]() {
      if (cur < 0) return;
      if (cur_slot == slot0 && cur >= b0 && cur - b0 + 1 < kDepSeg) {
        const int off = (int)(cur - b0);
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          atomicAdd(&seg[((o >> 1) * kDepSeg + off + (o & 1)) * 2 + cur_s], W[o]);
          W[o] = 0;
        }
        return;
      }
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int64_t node = cur + (o & 1) + X * ((o >> 1) & 1) + XY * (o >> 2);
        atomicAdd(&A.Nfix[2 * node + cur_s], W[o]);
        W[o] = 0;
      }
    };
    for (int64_t q = q0; q < q1; ++q) {
      const int sl = p.sim[q];
      const SimDev &sd = A.sims[sl];
      const int i = p.c[0][q], j = p.c[1][q], k = p.c[2][q];
      if (i < 0 || j < 0 || k < 0 || i > sd.nx || j > sd.ny || k > sd.nz) continue;
      const int64_t Xs = sd.X, XYs = (int64_t)sd.X * sd.Y;
      const int64_t node = sd.gbase + i + Xs * j + XYs * k;
      const int s = p.sp[q] - 1;
      if (node != cur || s != cur_s) {
        flush();
        cur = node;
        cur_s = s;
        cur_slot = sl;
        X = Xs;
        XY = XYs;
      }
      const float fx = p.f[0][q], fy = p.f[1][q], fz = p.f[2][q];
      const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const float w = wx[o & 1] * wy[(o >> 1) & 1] * wz[o >> 2];
        W[o] += (unsigned long long)(double(w) * kFix);
      }
    }
    flush();
    __syncthreads();
    if (slot0 >= 0) {
      const SimDev &s0 = A.sims[slot0];
      const int64_t X0 = s0.X, XY0 = (int64_t)s0.X * s0.Y;
      for (int e = threadIdx.x; e < 4 * kDepSeg * 2; e += kT) {
        const unsigned long long v = seg[e];
        if (!v) continue;
        const int sidx = e & 1, r = e >> 1;
        const int sg = r / kDepSeg, off = r % kDepSeg;
        const int64_t node = b0 + off + X0 * (sg & 1) + XY0 * (sg >> 1);
        atomicAdd(&A.Nfix[2 * node + sidx], v);
      }
    }
    __syncthreads();                    // the segments are cleared for the next chunk
  }
}


// ---------------------------------------------------------------------------
// per-node kernels over node tiles {slot, first node}: 256 nodes each

__device__ __forceinline__ void node_ijk(const SimDev &sd, int64_t node, int &i, int &j, int &k) {
  i = (int)(node % sd.X) + 1;
  j = (int)((node / sd.X) % sd.Y) + 1;
  k = (int)(node / ((int64_t)sd.X * sd.Y)) + 1;
}

// guard folds (ParticlesGridify.f90:129-285): x flow, y and z periodic
__device__ __forceinline__ int fold_src(int m, int n, bool flow, int src[2], bool *zero) {
  *zero = false;
  if (flow) {
    if (m == 1 || m == n + 2) { *zero = true; return 0; }
    if (m == 2) { src[0] = 2; src[1] = 1; return 2; }
    if (m == n + 1) { src[0] = n + 1; src[1] = n + 2; return 2; }
  } else {
    if (m == 1) { src[0] = 1; src[1] = n + 1; return 2; }
    if (m == 2) { src[0] = 2; src[1] = n + 2; return 2; }
    if (m == n + 1) { src[0] = n + 1; src[1] = 1; return 2; }
    if (m == n + 2) { src[0] = n + 2; src[1] = 2; return 2; }
  }
  src[0] = m;
  return 1;
}

// N = W 2^-40 n2p (integer folds first), C = sum_s N charge, and the
// potential buffers cleared (BCpotentials.f90:36).
__global__ void __launch_bounds__(kT) k_fold(Arenas A) {
  const int nt = A.g->n_node_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.node_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    const int64_t ln = (int64_t)tl.a + threadIdx.x;
    if (ln >= sd.G) continue;
    int i, j, k;
    node_ijk(sd, ln, i, j, k);
    int sx[2], sy[2], sz[2];
    bool zx, zy, zz;
    const int nxs = fold_src(i, sd.nx, true, sx, &zx);
    const int nys = fold_src(j, sd.ny, false, sy, &zy);
    const int nzs = fold_src(k, sd.nz, false, sz, &zz);
    const int64_t node = sd.gbase + ln;
    double cacc = 0.0;
    for (int s = 0; s < 2; ++s) {
      unsigned long long w = 0;
      if (!zx && !zy && !zz)
        for (int a = 0; a < nxs; ++a)
          for (int b = 0; b < nys; ++b)
            for (int c = 0; c < nzs; ++c)
              w += A.Nfix[2 * (sd.gbase + (sx[a] - 1) + (int64_t)sd.X * ((sy[b] - 1) +
                                                           (int64_t)sd.Y * (sz[c] - 1))) + s];
      const double v = (double)w * kFixInv * sd.n2p[s];
      A.N[2 * node + s] = v;
      cacc = cacc + v * sd.charge[s];
    }
    A.C[node] = cacc;
    A.U[node] = 0.0;
    A.u[0][node] = 0.0;
    A.u[1][node] = 0.0;
  }
}

// |b|^2 (b = -C / (dV eps0), CalcEpotential.f90:50) per live slot, one block
// each, in a fixed order; the slot's stop threshold and solve state.
__global__ void __launch_bounds__(kT) k_bnorm(Arenas A) {
  __shared__ double sh[kT / 64];
  const int nl = A.g->n_live;
  for (int q = blockIdx.x; q < nl; q += gridDim.x) {
    const int s = A.live[q];
    const SimDev &sd = A.sims[s];
    double acc = 0.0;
    const int64_t X = sd.X, XY = (int64_t)sd.X * sd.Y;
    const uint32_t nx = (uint32_t)sd.nx, ny = (uint32_t)sd.ny, cells = (uint32_t)sd.cells;
    for (uint32_t c = threadIdx.x; c < cells; c += kT) {
      const uint32_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
      const double b = -A.C[sd.gbase + (x + 1) + X * (y + 1) + XY * (z + 1)] / sd.dVeps0;
      acc = acc + b * b;
    }
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) {
      SimState &st = A.state[s];
      const double th = sd.maxerr * sqrt(tot);
      st.thresh2 = th * th;
      st.ifixq = st.thresh2 > 0.0 ? (double)kStopUnits / st.thresh2 : 1e300;
      st.j = 0;
      st.base = 0;
      st.solving = 1;
      st.noconv = 0;
      st.have_e = 0;
      st.ucur = 0;
      st.stop_j = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// BCpotentials.f90:51-77 on the two x faces: U(face node) = 1/(4 pi eps0)
// sum over cells C * D(offset), D of CalcDistance.f90:25-91 for x flow with y
// and z periodic.  Face tiles {slot, face, first item, x-plane chunk}: an
// item is kFaceNodes face nodes consecutive in y (one y row), one item per
// thread; partial sums per chunk, summed in chunk order afterwards.
constexpr int kFaceNodes = 8;

__global__ void __launch_bounds__(kT) k_distance(Arenas A, int slot) {
  const SimDev &sd = A.sims[slot];
  const int64_t na = sd.nx + 1, nb = sd.ny + 1, nc = sd.nz + 1;
  const int64_t total = na * nb * nc;
  const double d = sd.dx;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kT) {
    const int b = (int)(t % nb), c = (int)((t / nb) % nc), a = (int)(t / (nb * nc));
    const double di = (double)a, dj = (double)b, dk = (double)c;
    const double Ny = (double)sd.ny, Nz = (double)sd.nz;
    double temp = 0.0;
    if (a != 0 || b != 0 || c != 0)
      temp = 1.0 / sqrt((di * d) * (di * d) + (dj * d) * (dj * d) + (dk * d) * (dk * d));
    temp = temp + 1.0 / sqrt((di * d) * (di * d) + ((Ny - dj) * d) * ((Ny - dj) * d) +
                             (dk * d) * (dk * d));
    temp = temp + 1.0 / sqrt((di * d) * (di * d) + (dj * d) * (dj * d) +
                             ((Nz - dk) * d) * ((Nz - dk) * d));
    temp = temp + 1.0 / sqrt((di * d) * (di * d) + ((Ny - dj) * d) * ((Ny - dj) * d) +
                             ((Nz - dk) * d) * ((Nz - dk) * d));
    A.D[sd.dbase + t] = temp;
  }
}

// Each node's sum runs over x planes, then z rows, then y, in that order, one
// FMA per cell.  For a block of kFaceNodes y cells jb..jb+R-1 of a row, the
// item's nodes jj0..jj0+R-1 need D at |jb - jj0 + u - (R - 1)|, u < 2R - 1: a
// window of the row that slides by R per block, so a thread loads R values of
// D and R of C (the same for the whole wave) per R*R FMAs.
constexpr int kFaceRowMax = 512;      // y rows staged in LDS (longer rows read C from memory)

// The FMAs of one (x plane, z row) of cells for the thread's kFaceNodes nodes,
// C of the row from `crow` (LDS or memory, element stride cs).
template <typename CRow>
__device__ __forceinline__ void facesum_row(double (&acc)[kFaceNodes], const double *Drow, CRow crow,
                                            int ny, int jj0) {
  constexpr int R = kFaceNodes;
  // D index of window slot u for the block starting at jb (clamped into the
  // row: a clamped slot only meets a node or cell beyond the row)
  auto widx = [&](int jb, int u) {
    int d = jb - jj0 + u - (R - 1);
    d = d < 0 ? -d : d;
    return d < ny ? d : ny;
  };
  double w[2 * R - 1];
#pragma unroll
  for (int u = 0; u < 2 * R - 1; ++u) w[u] = Drow[widx(0, u)];
  for (int jb = 0; jb < ny; jb += R) {
    if (jb + R <= ny) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double c = crow(jb + q);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(c, w[q - r + R - 1], acc[r]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (jb + q < ny) {
          const double c = crow(jb + q);
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(c, w[q - r + R - 1], acc[r]);
        }
      }
    }
    if (jb + R < ny) {
#pragma unroll
      for (int u = 0; u < R - 1; ++u) w[u] = w[u + R];
#pragma unroll
      for (int u = R - 1; u < 2 * R - 1; ++u) w[u] = Drow[widx(jb + R, u)];
    }
  }
}

// All threads of a tile walk the same (x plane, z row) sequence, so each row
// of C is loaded once per block into LDS (double-buffered: one barrier per
// row) and read from there by every wave.
__global__ void __launch_bounds__(kT) k_facesum(Arenas A) {
  constexpr int R = kFaceNodes;
  __shared__ double cs[2][kFaceRowMax];
  const int nt = A.g->n_face_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.face_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    const int face = tl.a, chunk = tl.c;
    const int ny = sd.ny, nz = sd.nz;
    const int gy = (ny + R - 1) / R;                  // items per y row
    const int item = tl.b + (int)threadIdx.x;
    const bool on = item < gy * nz;
    const int jj0 = on ? (item % gy) * R : 0, kk = on ? item / gy : 0;
    const int i0 = chunk * kFacePlanes;
    const int i1 = i0 + kFacePlanes < sd.nx ? i0 + kFacePlanes : sd.nx;
    const int64_t X = sd.X, XY = (int64_t)sd.X * sd.Y;
    const int64_t nb = ny + 1, nbc = nb * (nz + 1);
    const double *Dt = A.D + sd.dbase;
    const double *C0 = A.C + sd.gbase;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // rows in order: it = (ip - i0) nz + kp
    auto crow_ptr = [&](int it) {
      const int ip = i0 + it / nz, kp = it % nz;
      return C0 + (ip + 1) + X * 1 + XY * (kp + 1);
    };
    auto drow_ptr = [&](int it) {
      const int ip = i0 + it / nz, kp = it % nz;
      const int a = face == 0 ? ip + 1 : sd.nx - ip;
      const int dk = kp > kk ? kp - kk : kk - kp;
      return Dt + a * nbc + dk * nb;
    };
    const int rows = (i1 - i0) * nz;
    if (ny <= kFaceRowMax) {                          // block-uniform
      {
        const double *Crow = crow_ptr(0);
        for (int q = threadIdx.x; q < ny; q += kT) cs[0][q] = Crow[X * q];
      }
      __syncthreads();
      for (int it = 0; it < rows; ++it) {
        if (it + 1 < rows) {
          const double *Crow = crow_ptr(it + 1);
          for (int q = threadIdx.x; q < ny; q += kT) cs[(it + 1) & 1][q] = Crow[X * q];
        }
        const double *cb = cs[it & 1];
        facesum_row(acc, drow_ptr(it), [&](int q) { return cb[q]; }, ny, jj0);
        __syncthreads();
      }
    } else {
      for (int it = 0; it < rows; ++it) {
        const double *Crow = crow_ptr(it);
        facesum_row(acc, drow_ptr(it), [&](int q) { return Crow[X * q]; }, ny, jj0);
      }
    }
    if (on) {
      const int64_t fnodes = (int64_t)ny * nz;
      double *out = A.fscr + sd.fbase + ((int64_t)face * sd.fchunks + chunk) * fnodes +
                    (int64_t)ny * kk;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (jj0 + r < ny) out[jj0 + r] = acc[r];
    }
  }
}

// Partials summed in chunk order; the face values go to both solve buffers.
__global__ void __launch_bounds__(kT) k_facereduce(Arenas A) {
  const int nt = A.g->n_fred_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.fred_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    const int face = tl.a;
    const int64_t fnodes = (int64_t)sd.ny * sd.nz;
    const int64_t fn = (int64_t)tl.b + threadIdx.x;
    if (fn >= fnodes) continue;
    double s = 0.0;
    for (int c = 0; c < sd.fchunks; ++c)
      s = s + A.fscr[sd.fbase + ((int64_t)face * sd.fchunks + c) * fnodes + fn];
    const int jj = (int)(fn % sd.ny), kk = (int)(fn / sd.ny);
    const int i = face == 0 ? 0 : sd.nx + 1;
    const int64_t node = sd.gbase + i + (int64_t)sd.X * (jj + 1) + (int64_t)sd.X * sd.Y * (kk + 1);
    const double v = s * sd.kappa;
    A.u[0][node] = v;
    A.u[1][node] = v;
  }
}

// ---------------------------------------------------------------------------
// the mixed-precision solve

// Refinement pass over cell tiles: u' = u + e_m (the previous refinement's
// last correction, if any) into the other FP64 buffer, r = b - A u' in FP64
// with fresh periodic images, r32 = r.  The block completing a slot's tiles
// folds the inner count into the base and flips the buffers.
__global__ void __launch_bounds__(kT) k_outer(Arenas A, int cpt) {
  const int nt = A.g->n_cell_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.cell_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    SimState &st = A.state[tl.slot];
    if (!st.solving) continue;
    const bool upd = st.have_e != 0;
    const double *uc = A.u[st.ucur] + sd.gbase;
    double *un = A.u[st.ucur ^ 1] + sd.gbase;
    const float *e = A.e32[st.j % kEBufs] + sd.cbase;
    const int64_t X = sd.X, XY = (int64_t)sd.X * sd.Y;
    const uint32_t nx = (uint32_t)sd.nx, ny = (uint32_t)sd.ny, nz = (uint32_t)sd.nz;
    const uint32_t pl = nx * ny, cells = (uint32_t)sd.cells;
    // the slot's constants in registers before the first store
    const double i2 = sd.i2, dVeps0 = sd.dVeps0;
    const double *Cs = A.C + sd.gbase;
    float *r32 = A.r32 + sd.cbase;
    const uint32_t c0 = (uint32_t)tl.a + threadIdx.x;
    for (int q = 0; q < cpt; ++q) {
      const uint32_t c = c0 + q * kT;
      if (c >= cells) break;
      const uint32_t x = c % nx, y = (c / nx) % ny, z = c / pl;
      const int64_t node = (x + 1) + X * (y + 1) + XY * (z + 1);
      const uint32_t yn = y + 1 < ny ? y + 1 : 0, ys = y > 0 ? y - 1 : ny - 1;
      const uint32_t zt = z + 1 < nz ? z + 1 : 0, zb = z > 0 ? z - 1 : nz - 1;
      const int64_t nN = (x + 1) + X * (yn + 1) + XY * (z + 1), nS = (x + 1) + X * (ys + 1) + XY * (z + 1);
      const int64_t nT = (x + 1) + X * (y + 1) + XY * (zt + 1), nB = (x + 1) + X * (y + 1) + XY * (zb + 1);
      const uint32_t cE = c + 1, cW = c - 1;
      const uint32_t cN = x + nx * yn + pl * z, cS = x + nx * ys + pl * z;
      const uint32_t cT = x + nx * y + pl * zt, cB = x + nx * y + pl * zb;
      double uC = uc[node], uE = uc[node + 1], uW = uc[node - 1];
      double uN = uc[nN], uS = uc[nS], uT = uc[nT], uB = uc[nB];
      if (upd) {
        uC += (double)e[c];
        if (x + 1 < nx) uE += (double)e[cE];
        if (x > 0) uW += (double)e[cW];
        uN += (double)e[cN];
        uS += (double)e[cS];
        uT += (double)e[cT];
        uB += (double)e[cB];
        un[node] = uC;
      }
      const double b = -Cs[node] / dVeps0;
      const double lap = ((uE + uW - 2.0 * uC) + (uN + uS - 2.0 * uC) + (uT + uB - 2.0 * uC)) * i2;
      r32[c] = (float)(b - lap);
    }
  }
}

// After a refinement pass, per solving slot: the inner count folds into the
// base and the FP64 buffers flip.
__global__ void k_outer_done(Arenas A) {
  const int nl = A.g->n_live;
  if (blockIdx.x == 0 && threadIdx.x == 0) A.g->chunk_j0 = 0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nl; q += gridDim.x * blockDim.x) {
    SimState &st = A.state[A.live[q]];
    if (!st.solving) continue;
    if (st.have_e) {
      st.ucur ^= 1;
      st.base += st.j;
    }
    st.have_e = 1;
    st.j = 0;
  }
}

// One FP32 Jacobi sweep e_{j+1} = J(e_j) on every solving slot's cell tiles,
// with the stop test of iterate base + j from the same neighbour sums.
// base[i] through a 32-bit byte offset (a slot holds fewer than 2^30 cells),
// so the access takes the uniform base and one 32-bit offset register.
__device__ __forceinline__ float ldf(const float *base, uint32_t i) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (i << 2));
}

__device__ __forceinline__ void stf(float *base, uint32_t i, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (i << 2)) = v;
}

// v / d for v < 2^22 (one float product, one fix-up).
__device__ __forceinline__ uint32_t small_quot(uint32_t v, uint32_t d, float invd) {
  uint32_t q = (uint32_t)((float)v * invd);
  const int32_t r = (int32_t)v - (int32_t)(q * d);
  if (r < 0) --q;
  else if (r >= (int32_t)d) ++q;
  return q;
}

// Residual units min(u2, 2^40 + 1) truncated to an integer: for a value below
// 2^52, trunc(v) + 2^52 holds the integer in its low mantissa bits (NaN and
// infinities clamp to the cap, as the comparison with the cap does).
__device__ __forceinline__ unsigned long long units_of(double u2) {
  const double v = __builtin_trunc(__builtin_fmin(u2, (double)(kStopUnits + 1))) + 4503599627370496.0;
  return (unsigned long long)__double_as_longlong(v) - 0x4330000000000000ull;
}

__device__ __forceinline__ float ldb(const float *base, uint32_t byte_off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byte_off);
}

// The operands of a thread's cells c0 + q kT (coordinates x, y, z of the
// first, stepped by (sx, sy, sz) per kT cells) for an iterate j > 0: r, e at
// the cell and its six neighbours (periodic in y and z, zero beyond the x
// faces), through byte offsets from the uniform bases; all loads before the
// first sum.  FULL: every cell inside the slot; otherwise cells beyond it read
// the slot's last cell (values unused).
template <int CPT, bool FULL>
__device__ __forceinline__ void sweep_gather(const float *ein, const float *r, uint32_t cells,
                                             uint32_t nx, uint32_t ny, uint32_t nz, uint32_t x,
                                             uint32_t y, uint32_t z, uint32_t c0, uint32_t sx,
                                             uint32_t sy, uint32_t sz, float (&rr)[CPT],
                                             float (&ec)[CPT], float (&sum6)[CPT]) {
  const uint32_t pl = nx * ny;
  const uint32_t bnx = nx * 4u, bpl = pl * 4u, bwy = (ny - 1) * bnx, bwz = (nz - 1) * bpl;
  float vE[CPT], vW[CPT], vN[CPT], vS[CPT], vT[CPT], vB[CPT];
  bool hasE[CPT], hasW[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    if (q > 0) {
      x += sx;
      y += sy;
      z += sz;
      if (x >= nx) {
        x -= nx;
        ++y;
      }
      if (y >= ny) {
        y -= ny;
        ++z;
      }
    }
    const bool in = FULL || z < nz;
    const uint32_t xl = in ? x : nx - 1, yl = in ? y : ny - 1, zl = in ? z : nz - 1;
    const uint32_t cb = (in ? c0 + q * kT : cells - 1) * 4u;
    hasE[q] = xl + 1 < nx;
    hasW[q] = xl > 0;
    rr[q] = ldb(r, cb);
    ec[q] = ldb(ein, cb);
    vE[q] = ldb(ein, hasE[q] ? cb + 4u : cb);
    vW[q] = ldb(ein, hasW[q] ? cb - 4u : cb);
    vN[q] = ldb(ein, yl + 1 < ny ? cb + bnx : cb - bwy);
    vS[q] = ldb(ein, yl > 0 ? cb - bnx : cb + bwy);
    vT[q] = ldb(ein, zl + 1 < nz ? cb + bpl : cb - bwz);
    vB[q] = ldb(ein, zl > 0 ? cb - bpl : cb + bwz);
  }
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const float eE = hasE[q] ? vE[q] : 0.0f;
    const float eW = hasW[q] ? vW[q] : 0.0f;
    sum6[q] = ((eE + eW) + (vN[q] + vS[q])) + (vT[q] + vB[q]);
  }
}

template <int CPT, bool TUNE>
__global__ void __launch_bounds__(kT) k_sweep(Arenas A, int inner, int jj) {
  __shared__ unsigned long long wsum[kT / 64];
  const int nt = A.g->n_cell_tiles;
  // every solving simulation is at the same inner sweep: jj-th of the chunk
  const int j = TUNE ? 1 : A.g->chunk_j0 + jj;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.cell_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    SimState &st = A.state[tl.slot];
    if (!TUNE && (!st.solving || j >= inner)) continue;
    const float *ein = A.e32[j % kEBufs] + sd.cbase;
    float *eout = A.e32[(j + 1) % kEBufs] + sd.cbase;
    const float *r = A.r32 + sd.cbase;
    const double ifixq = st.ifixq;
    // the slot's constants in registers before the first store (a store
    // through eout could otherwise alias them and force a reload per cell)
    const float wf = sd.wf, kf = sd.kf, aff = sd.aff, i2f = sd.i2f;
    const int64_t ubase = sd.ubase;
    unsigned long long units = 0;
    // the thread's cells a + q kT + tid, q < CPT: every load and store of a
    // wave is one contiguous run (32-bit index arithmetic: a simulation holds
    // fewer than 2^31 cells).  The kernel is bound by instruction issue, so
    // the cell coordinates are not divided out per cell: the tile's first
    // cell is split once (uniform), the thread's offset carried in with two
    // small quotients, and each further cell steps by kT cells with two
    // carries.
    const uint32_t nx = (uint32_t)sd.nx, ny = (uint32_t)sd.ny, nz = (uint32_t)sd.nz;
    const uint32_t cells = (uint32_t)sd.cells, pl = nx * ny;
    const uint32_t a = (uint32_t)tl.a;
    const uint32_t za = a / pl, ra = a - za * pl, ya = ra / nx, xa = ra - ya * nx;
    const uint32_t kq = (uint32_t)kT / nx, sx = (uint32_t)kT - kq * nx, sy = kq % ny, sz = kq / ny;
    uint32_t x = xa + threadIdx.x, y = ya, z = za;
    {
      const uint32_t dy = small_quot(x, nx, 1.0f / (float)nx);
      x -= dy * nx;
      y += dy;
      const uint32_t dz = small_quot(y, ny, 1.0f / (float)ny);
      y -= dz * ny;
      z += dz;
    }
    const uint32_t c0 = a + threadIdx.x;
    // every load unconditional (cells beyond the slot read its last cell,
    // values unused), all of them before the first sum, so a thread waits
    // for memory once per tile
    float rr[CPT], ec[CPT], sum6[CPT];
    if (j > 0) {                                       // uniform
      if (a + (uint32_t)(CPT * kT) <= cells)           // every tile but a slot's last
        sweep_gather<CPT, true>(ein, r, cells, nx, ny, nz, x, y, z, c0, sx, sy, sz, rr, ec, sum6);
      else
        sweep_gather<CPT, false>(ein, r, cells, nx, ny, nz, x, y, z, c0, sx, sy, sz, rr, ec, sum6);
    } else {                                           // e_0 = 0
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const uint32_t c = c0 + q * kT;
        rr[q] = ldf(r, c < cells ? c : cells - 1);
        ec[q] = 0.0f;
        sum6[q] = 0.0f;
      }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const uint32_t c = c0 + q * kT;
      if (c < cells) {
        stf(eout, c, __builtin_fmaf(wf, sum6[q], __builtin_fmaf(kf, ec[q], -aff * rr[q])));
        const float res = rr[q] - i2f * (sum6[q] - 6.0f * ec[q]);
        const double u2 = ((double)res * (double)res) * ifixq;
        units += units_of(u2);
      }
    }
    // wave, then block sum of the integer units, stored for the tile
    for (int off = 32; off > 0; off >>= 1) units += __shfl_down(units, off, 64);
    if (lane_id() == 0) wsum[threadIdx.x >> 6] = units;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long sum = 0;
      for (int w = 0; w < kT / 64; ++w) sum += wsum[w];
      A.units[A.ucap * (j % kDecideEvery) + ubase + tl.b] = sum;
    }
    __syncthreads();
  }
}

// After kDecideEvery sweeps (the first of them the jd-th of the chunk), per
// solving slot (one block each): the residuals of iterates base + j .. base +
// j + kDecideEvery - 1 from their tiles' units, and the stop rule applied to
// each in turn, as after every sweep.
__global__ void __launch_bounds__(kT) k_decide(Arenas A, int inner, int jd) {
  __shared__ unsigned long long wsum[kDecideEvery][kT / 64];
  const int nl = A.g->n_live;
  const int j0 = A.g->chunk_j0 + jd;
  for (int q = blockIdx.x; q < nl; q += gridDim.x) {
    const int s = A.live[q];
    const SimDev &sd = A.sims[s];
    SimState &st = A.state[s];
    if (!st.solving || j0 >= inner) continue;         // block-uniform
    // the group's iterates summed in one pass over the tiles
    const int tiles = sd.tiles;
    const unsigned long long *u = A.units + sd.ubase;
    unsigned long long acc[kDecideEvery];
#pragma unroll
    for (int k = 0; k < kDecideEvery; ++k) acc[k] = 0;
    for (int t = threadIdx.x; t < tiles; t += kT) {
#pragma unroll
      for (int k = 0; k < kDecideEvery; ++k) acc[k] += u[A.ucap * ((j0 + k) % kDecideEvery) + t];
    }
#pragma unroll
    for (int k = 0; k < kDecideEvery; ++k) {
      for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_down(acc[k], off, 64);
      if (lane_id() == 0) wsum[k][threadIdx.x >> 6] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int64_t cap = 2 * (int64_t)sd.maxit + 1;
      for (int k = 0; k < kDecideEvery; ++k) {
        unsigned long long tot = 0;
        for (int w = 0; w < kT / 64; ++w) tot += wsum[k][w];
        const int j = j0 + k;
        const int64_t it = st.base + j;
        if (tot <= kStopUnits) {                 // |res| <= maxerr |b| (CalcEpotential.f90:60-62)
          st.solving = 0;
          st.sweeps = it;
          st.stop_j = j;
          break;
        }
        if (it + 1 >= cap) {                     // 2*maxit+1 sweeps (:67-69)
          st.solving = 0;
          st.sweeps = cap;
          st.stop_j = j + 1;
          st.noconv = 1;
          break;
        }
        st.j = j + 1;
      }
    }
    __syncthreads();
  }
}

// Simulations still solving, to the device counter and the host word.
__global__ void k_count_active(Arenas A, int32_t *host_active) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int nl = A.g->n_live;
  for (int q = threadIdx.x; q < nl; q += blockDim.x)
    if (A.state[A.live[q]].solving) atomicAdd(&cnt, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    A.g->active = cnt;
    A.g->chunk_j0 += kChunkSweeps;
    *reinterpret_cast<volatile int32_t *>(host_active) = cnt;
  }
}

// Final potential: u + e at the stop iterate on the interior, the face values
// on the x guards, periodic images on the y and z guards (CalcEpotential.f90:
// 100-239 halo order z, y, x: the x-guard planes' y/z guard rows are images
// of the face values).
__global__ void __launch_bounds__(kT) k_finish(Arenas A) {
  const int nt = A.g->n_node_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.node_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    const SimState &st = A.state[tl.slot];
    const int64_t ln = (int64_t)tl.a + threadIdx.x;
    if (ln >= sd.G) continue;
    int i, j, k;
    node_ijk(sd, ln, i, j, k);
    const int sj = j == 1 ? sd.ny + 1 : (j == sd.ny + 2 ? 2 : j);
    const int sk = k == 1 ? sd.nz + 1 : (k == sd.nz + 2 ? 2 : k);
    const int64_t src = (i - 1) + (int64_t)sd.X * ((sj - 1) + (int64_t)sd.Y * (sk - 1));
    double v = A.u[st.ucur][sd.gbase + src];
    if (i >= 2 && i <= sd.nx + 1 && st.stop_j > 0)
      v += (double)A.e32[st.stop_j % kEBufs][sd.cbase + (i - 2) +
                                        (int64_t)sd.nx * ((sj - 2) + (int64_t)sd.ny * (sk - 2))];
    A.U[sd.gbase + ln] = v;
  }
}

// E = -grad U in FP32 (CalcEfield.f90:27-34), zero on the x guard planes
// (BCfields.f90:31-90), periodic images on the y and z guards.
__global__ void __launch_bounds__(kT) k_efield(Arenas A) {
  const int nt = A.g->n_node_tiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const Tile tl = A.node_tiles[t];
    const SimDev &sd = A.sims[tl.slot];
    const int64_t ln = (int64_t)tl.a + threadIdx.x;
    if (ln >= sd.G) continue;
    int i, j, k;
    node_ijk(sd, ln, i, j, k);
    const int sj = j == 1 ? sd.ny + 1 : (j == sd.ny + 2 ? 2 : j);
    const int sk = k == 1 ? sd.nz + 1 : (k == sd.nz + 2 ? 2 : k);
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (i >= 2 && i <= sd.nx + 1) {
      const int64_t X = sd.X, XY = (int64_t)sd.X * sd.Y;
      const int64_t c = sd.gbase + (i - 1) + X * (sj - 1) + XY * (sk - 1);
      const double *U = A.U;
      ez = -(U[c + XY] - U[c - XY]) * 0.5 / sd.dx;
      ey = -(U[c + X] - U[c - X]) * 0.5 / sd.dx;
      ex = -(U[c + 1] - U[c - 1]) * 0.5 / sd.dx;
    }
    A.E32[sd.gbase + ln] = make_float4((float)ex, (float)ey, (float)ez, 0.0f);
  }
}

// AdvanceVelocity.f90:39-121 in FP32 (the split form's push).
__global__ void __launch_bounds__(kT) k_boris(Arenas A, PartArena p) {
  const int64_t n = A.g->n;
  for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < n; q += (int64_t)gridDim.x * kT) {
    const SimDev &sd = A.sims[p.sim[q]];
    const int i = p.c[0][q], j = p.c[1][q], k = p.c[2][q];
    if (i < 0 || j < 0 || k < 0 || i > sd.nx || j > sd.ny || k > sd.nz) continue;
    const float fx = p.f[0][q], fy = p.f[1][q], fz = p.f[2][q];
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    const int64_t X = sd.X, XY = (int64_t)sd.X * sd.Y;
    const int64_t base = sd.gbase + i + X * j + XY * k;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const float w = wx[a] * wy[b] * wz[c];
          const float4 e = A.E32[base + a + X * b + XY * c];
          ex += e.x * w;
          ey += e.y * w;
          ez += e.z * w;
        }
    const int s = p.sp[q] - 1;
    const float h = sd.etadt2[s];
    const float *tt = sd.tB[s], *sb = sd.sB[s];
    const float vm1 = p.v[0][q] - sd.ve[0] + h * ex;
    const float vm2 = p.v[1][q] - sd.ve[1] + h * ey;
    const float vm3 = p.v[2][q] - sd.ve[2] + h * ez;
    const float vd1 = vm1 + vm2 * tt[2] - vm3 * tt[1];
    const float vd2 = vm2 + vm3 * tt[0] - vm1 * tt[2];
    const float vd3 = vm3 + vm1 * tt[1] - vm2 * tt[0];
    const float vp1 = vm1 + vd2 * sb[2] - vd3 * sb[1];
    const float vp2 = vm2 + vd3 * sb[0] - vd1 * sb[2];
    const float vp3 = vm3 + vd1 * sb[1] - vd2 * sb[0];
    p.v[0][q] = vp1 + sd.ve[0] + h * ex;
    p.v[1][q] = vp2 + sd.ve[1] + h * ey;
    p.v[2][q] = vp3 + sd.ve[2] + h * ez;
  }
}

// End of a step: per live slot, the step count and the particles pushed; the
// per-slot status and the arena count to the mapped host arrays.
__global__ void k_step_end(Arenas A, picard_sim_status *host_status, int64_t *host_n,
                           int32_t *host_err) {
  const Global &gl = *A.g;
  for (int q = threadIdx.x; q < gl.n_live; q += blockDim.x) {
    const int s = A.live[q];
    SimState &st = A.state[s];
    st.step += 1;
    st.pushed = st.count;
    picard_sim_status o;
    o.live = 1;
    o.noconv = st.noconv;
    o.steps = st.step;
    o.num_local = st.count;
    o.pushed = st.pushed;
    o.sweeps = st.sweeps;
    host_status[s] = o;
  }
  if (threadIdx.x == 0) {
    *host_n = gl.n;
    *host_err = gl.err;
  }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = v;
}
__global__ void k_fill_u64(unsigned long long *p, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0ull;
}
__global__ void k_fill_f32(float *p, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0.0f;
}

// widened E of one slot for field reads
__global__ void k_widen(const float4 *E, int64_t n, double *out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const float4 e = E[q];
    out[3 * q] = e.x;
    out[3 * q + 1] = e.y;
    out[3 * q + 2] = e.z;
  }
}

// one slot's particles as real_particles(8, n), int_particles(n)
__global__ void k_to_aos(const SimDev *sims, PartArena p, int slot, int64_t first, int64_t count,
                         double *rp, int32_t *sp) {
  const SimDev &sd = sims[slot];
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = first + q;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double xi = (double)p.c[d][s] + (double)p.f[d][s];
      rp[8 * q + d] = sd.lmin[d] + (xi - 0.5) * sd.dx;
      rp[8 * q + 3 + d] = (double)p.v[d][s];
    }
    rp[8 * q + 6] = 0.0;
    rp[8 * q + 7] = (double)p.id[s];
    sp[q] = p.sp[s];
  }
}

// ---------------------------------------------------------------------------
// host side

struct Range {
  int64_t off, len;
};

// First-fit allocator over [0, cap).
class Allocator {
 public:
  explicit Allocator(int64_t cap = 0) { reset(cap); }
  void reset(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  bool alloc(int64_t len, int64_t *off) {
    if (len <= 0) {
      *off = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        *off = it->first;
        const int64_t rest = it->second - len, at = it->first + len;
        free_.erase(it);
        if (rest > 0) free_[at] = rest;
        return true;
      }
    }
    return false;
  }
  void release(int64_t off, int64_t len) {
    if (len <= 0) return;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

struct SlotHost {
  bool live = false;
  bool draining = false;      // removed; its particles leave the arena at the next step's sort
  picard_sim_params p{};
  picard_sim_derived d{};
  Range g{}, c{}, dt{}, f{}, k{};
  int64_t init_particles = 0;
  int64_t appended_at = -1;   // arena position of the initial load until the first step
};

}  // namespace ens
}  // namespace picard

using namespace picard;
using namespace picard::ens;

struct picard_ensemble {
  picard_ensemble_config cfg{};
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  Arenas A{};
  PartArena P[2]{};
  int cur = 0;                       // arena holding the sorted particles
  int64_t cap = 0;                   // particle capacity
  int64_t max_cells = 0, max_nodes = 0, max_dist = 0, max_face = 0;
  int32_t *scan_tmp = nullptr;
  int64_t scan_tmp_len = 0;
  int32_t *host_active = nullptr;    // mapped
  picard_sim_status *host_status = nullptr;   // mapped, per slot
  int64_t *host_n = nullptr;         // mapped
  int32_t *host_err = nullptr;       // mapped
  std::vector<SlotHost> slots;
  Allocator ag, ac, ad, af, ak;
  bool dirty = true;                 // work lists need rebuilding
  picard_ensemble_tile tile{4, 256, 4, 0};
  int inner = 512;
  // graphs: pre/post per arena parity, solve chunk with and without the
  // refinement pass
  hipGraphExec_t g_pre[2] = {nullptr, nullptr}, g_post[2] = {nullptr, nullptr};
  hipGraphExec_t g_chunk[2] = {nullptr, nullptr};
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};   // ends of the last two solve chunks enqueued
  int64_t n_host = 0;                // particles in the arena after the last step or load
  // capacities of the device work lists (entries)
  int64_t cap_node_t = 0, cap_cell_t = 0, cap_face_t = 0, cap_fred_t = 0;
  std::vector<void *> allocs;
};

namespace {

int dmalloc(picard_ensemble *e, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  PICARD_HIP(hipMalloc(p, bytes));
  e->allocs.push_back(*p);
  return PICARD_OK;
}

template <typename T>
int dalloc_n(picard_ensemble *e, T **p, int64_t n) {
  return dmalloc(e, reinterpret_cast<void **>(p), sizeof(T) * (size_t)(n > 0 ? n : 1));
}

void free_graphs(picard_ensemble *e) {
  for (hipGraphExec_t *g : {&e->g_pre[0], &e->g_pre[1], &e->g_post[0], &e->g_post[1],
                            &e->g_chunk[0], &e->g_chunk[1]}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace

// picard.f90:333-466 for the ensemble's plasma (species 1 = e-, 2 = H+),
// with the same operand order as the single-domain host (picard_amd/host.py).
static int derive_sim(const picard_sim_params *p, picard_sim_derived *d) {
  if (!p || !d) {
    set_error("picard_sim_derive: NULL");
    return PICARD_ERR_ARG;
  }
  if (p->nx < 8 || p->ny < 8 || p->nz < 8 || p->nx > 4096 || p->ny > 4096 || p->nz > 4096 ||
      (int64_t)p->nx * p->ny * p->nz > ((int64_t)1 << 30) ||
      p->ppc < 1 || p->ppc > 1024 || !(p->dx > 0.0) || !(p->density > 0.0) ||
      !(p->kelvin >= 0.0) || !(p->dt_wpe > 0.0) || !(p->maxerr > 0.0) || p->maxit < 1) {
    set_error("picard_sim_derive: parameters out of range");
    return PICARD_ERR_ARG;
  }
  std::memset(d, 0, sizeof(*d));
  const int n[3] = {p->nx, p->ny, p->nz};
  double dx[3];
  for (int c = 0; c < 3; ++c) {
    d->lmin[c] = -0.5 * p->dx * n[c];
    d->lmax[c] = 0.5 * p->dx * n[c];
    dx[c] = (d->lmax[c] - d->lmin[c]) / (double)n[c];     // :344-360
  }
  d->dxyz = dx[0];
  d->dV = dx[0] * dx[1] * dx[2];
  d->charge[0] = -kEC;                                    // :371
  d->charge[1] = kEC;
  const double mass[2] = {kME, kMP};
  // density(2) * ppc / ppc (:373); density(1) = sum_{s>=2} -density(s)
  // charge(s) / charge(1) (:381); ppc(1) = ipp(2) epp(2) (:392)
  const double d2 = p->density * (double)p->ppc / (double)p->ppc;
  const double dens[2] = {0.0 + -d2 * d->charge[1] / d->charge[0], d2};
  const int ppc[2] = {p->ppc, p->ppc};
  for (int s = 0; s < 2; ++s) {
    d->n2p[s] = d->dV * dens[s] * (1.0 / (double)ppc[s]);  // :394-401
    d->eta[s] = d->charge[s] / mass[s];                   // :405
  }
  // mass-weighted bulk velocity (:408-413): only H+ carries ions
  const double w = (double)p->ppc * (d->n2p[1] * mass[1] + 1.0 * d->n2p[0] * mass[0] +
                                     0.0 * d->n2p[1] * mass[1]);
  const double v0[3] = {p->drift, 0.0, 0.0};
  for (int c = 0; c < 3; ++c) d->ve[c] = (0.0 + w * v0[c]) / (0.0 + w);
  d->omega_pe = std::sqrt(dens[0] * kEC * kEC / (kEPS0 * mass[0]));
  d->dt = p->dt_wpe / d->omega_pe;
  // |B| and |tB|^2 with the C library's pow, as the host derives them
  double (*volatile powp)(double, double) = std::pow;
  const double B4 = std::sqrt(((0.0 + powp(p->B[0], 2.0)) + powp(p->B[1], 2.0)) +
                              powp(p->B[2], 2.0));
  for (int s = 0; s < 2; ++s) {                           // :452-466 (gyration on)
    for (int c = 0; c < 3; ++c) d->tB[s][c] = 0.5 * d->dt * p->B[c] * d->eta[s];
    d->tB4[s] = 0.5 * d->dt * B4 * d->eta[s];
    for (int c = 0; c < 3; ++c) d->sB[s][c] = 2.0 * d->tB[s][c] / (1.0 + powp(d->tB4[s], 2.0));
  }
  return PICARD_OK;
}

static void slot_sizes(const picard_sim_params *p, int64_t *nodes, int64_t *cells, int64_t *dist,
                       int64_t *face, int64_t *keys, int32_t *fchunks) {
  const int64_t nx = p->nx, ny = p->ny, nz = p->nz;
  *nodes = (nx + 2) * (ny + 2) * (nz + 2);
  *cells = nx * ny * nz;
  *dist = (nx + 1) * (ny + 1) * (nz + 1);
  *fchunks = (int32_t)((nx + kFacePlanes - 1) / kFacePlanes);
  *face = 2 * (int64_t)*fchunks * ny * nz;
  *keys = 2 * (nx + 1) * (ny + 1) * (nz + 1);
}

static SimDev make_simdev(const SlotHost &h, int slot) {
  const picard_sim_params &p = h.p;
  const picard_sim_derived &d = h.d;
  SimDev s;
  std::memset(&s, 0, sizeof(s));
  s.nx = p.nx;
  s.ny = p.ny;
  s.nz = p.nz;
  s.live = h.live ? 1 : 0;
  s.X = p.nx + 2;
  s.Y = p.ny + 2;
  s.Z = p.nz + 2;
  s.ppc = p.ppc;
  s.G = (int64_t)s.X * s.Y * s.Z;
  s.gbase = h.g.off;
  s.cells = (int64_t)p.nx * p.ny * p.nz;
  s.cbase = h.c.off;
  s.dbase = h.dt.off;
  s.fbase = h.f.off;
  s.kbase = (int32_t)h.k.off;
  s.nkeys = (int32_t)h.k.len;
  s.fchunks = (p.nx + kFacePlanes - 1) / kFacePlanes;
  for (int c = 0; c < 3; ++c) s.lmin[c] = d.lmin[c];
  s.xmax = d.lmax[0];
  s.dx = d.dxyz;
  s.dVeps0 = d.dV * kEPS0;
  s.kappa = 1.0 / (2.0 * kTau * kEPS0);                   // (2 tau eps0)^-1, BCpotentials.f90:71
  for (int q = 0; q < 2; ++q) {
    s.n2p[q] = d.n2p[q];
    s.charge[q] = d.charge[q];
    s.etadt2[q] = (float)(0.5 * d.dt * d.eta[q]);
    for (int c = 0; c < 3; ++c) {
      s.tB[q][c] = (float)d.tB[q][c];
      s.sB[q][c] = (float)d.sB[q][c];
    }
  }
  for (int c = 0; c < 3; ++c) s.ve[c] = (float)d.ve[c];
  s.dtdx = (float)(d.dt / d.dxyz);
  s.i2 = 1.0 / (d.dxyz * d.dxyz);
  // cubic cells: ax = ay = az = 1/6, af = dx^2 / 6 (CalcEpotential.f90:44-47)
  s.wf = 1.0f / 6.0f;
  s.kf = (float)(1.0 - 6.0 * (double)s.wf);
  s.aff = (float)(d.dxyz * d.dxyz / 6.0);
  s.i2f = (float)s.i2;
  s.maxerr = p.maxerr;
  s.maxit = p.maxit;
  s.seed = p.seed;
  // sources: the plasma tables of the single-domain form (species 1 = e-, 2 = H+)
  PlasmaConst &pc = s.pc;
  pc.species = 2;
  pc.first_species = 1;
  pc.ipp[1] = p.ppc;
  pc.epp[1] = 1;
  pc.group[0] = 1;
  pc.group[1] = 2;
  pc.per_cell = 2 * p.ppc;
  pc.mass[0] = kME;
  pc.mass[1] = kMP;
  pc.kelvin[0] = pc.kelvin[1] = p.kelvin;
  for (int q = 0; q < 2; ++q) {
    pc.v0[q][0] = p.drift;
    for (int c = 0; c < 3; ++c) pc.tB[q][c] = d.tB[q][c];
  }
  for (int c = 0; c < 3; ++c) pc.ve[c] = d.ve[c];
  s.inflow = 2 * (int64_t)p.ny * p.nz * 2 * p.ppc;
  (void)slot;
  return s;
}

// Rebuild the device tables and work lists after a membership or tile change.
static int rebuild(picard_ensemble *e) {
  const int S = (int)e->slots.size();
  std::vector<SimDev> sims(S);
  std::vector<Tile> node_t, cell_t, face_t, fred_t, imap;
  std::vector<int32_t> live;
  const int64_t tcells = (int64_t)e->tile.threads * e->tile.cells_per_thread;
  int64_t pfirst = 0, ions = 0, ubase = 0;
  for (int s = 0; s < S; ++s) {
    const SlotHost &h = e->slots[s];
    sims[s] = make_simdev(h, s);
    if (!h.live) continue;
    live.push_back(s);
    SimDev &sd = sims[s];
    for (int64_t a = 0; a < sd.G; a += kT) node_t.push_back({s, (int32_t)a, 0, 0});
    // cell tiles: each slot's tiles cut into 8 contiguous shares (one per
    // XCD) and interleaved, so that with a launch of a multiple of 8 blocks
    // block b (on XCD b % 8) works on one share of every slot
    const int64_t nt = (sd.cells + tcells - 1) / tcells;
    sd.tiles = (int32_t)nt;
    sd.ubase = ubase;
    ubase += nt;
    std::vector<std::vector<Tile>> share(8);
    for (int x = 0; x < 8; ++x) {
      const int64_t t0 = nt * x / 8, t1 = nt * (x + 1) / 8;
      for (int64_t t = t0; t < t1; ++t)
        share[x].push_back({s, (int32_t)(t * tcells), (int32_t)t, 0});
    }
    size_t mx = 0;
    for (auto &v : share) mx = std::max(mx, v.size());
    for (size_t m = 0; m < mx; ++m)
      for (int x = 0; x < 8; ++x)
        if (m < share[x].size()) cell_t.push_back(share[x][m]);
    const int64_t fn = (int64_t)sd.ny * sd.nz;
    const int64_t items = (int64_t)((sd.ny + kFaceNodes - 1) / kFaceNodes) * sd.nz;
    for (int face = 0; face < 2; ++face) {
      for (int64_t a = 0; a < items; a += kT)
        for (int ch = 0; ch < sd.fchunks; ++ch) face_t.push_back({s, face, (int32_t)a, ch});
      for (int64_t a = 0; a < fn; a += kT) fred_t.push_back({s, face, (int32_t)a, 0});
    }
    sd.pfirst = pfirst;
    imap.push_back({s, (int32_t)ions, 0, 0});
    pfirst += sd.inflow;
    ions += sd.inflow / 2;
  }
  if ((int64_t)node_t.size() > e->cap_node_t || (int64_t)cell_t.size() > e->cap_cell_t ||
      (int64_t)face_t.size() > e->cap_face_t || (int64_t)fred_t.size() > e->cap_fred_t ||
      ions > INT32_MAX) {
    set_error("ensemble: work lists exceed their capacity");
    return PICARD_ERR_CAPACITY;
  }
  Global gl{};
  PICARD_HIP(hipMemcpyAsync(&gl, e->A.g, sizeof(Global), hipMemcpyDeviceToHost, e->stream));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  gl.inflow_total = pfirst;
  gl.n_node_tiles = (int32_t)node_t.size();
  gl.n_cell_tiles = (int32_t)cell_t.size();
  gl.n_face_tiles = (int32_t)face_t.size();
  gl.n_fred_tiles = (int32_t)fred_t.size();
  gl.n_live = (int32_t)live.size();
  gl.n_inflow_map = (int32_t)imap.size();
  auto up = [&](void *dst, const void *src, size_t bytes) -> int {
    if (bytes) PICARD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
    return PICARD_OK;
  };
  PICARD_TRY(up(e->A.sims, sims.data(), sizeof(SimDev) * S));
  PICARD_TRY(up(e->A.node_tiles, node_t.data(), sizeof(Tile) * node_t.size()));
  PICARD_TRY(up(e->A.cell_tiles, cell_t.data(), sizeof(Tile) * cell_t.size()));
  PICARD_TRY(up(e->A.face_tiles, face_t.data(), sizeof(Tile) * face_t.size()));
  PICARD_TRY(up(e->A.fred_tiles, fred_t.data(), sizeof(Tile) * fred_t.size()));
  PICARD_TRY(up(e->A.live, live.data(), sizeof(int32_t) * live.size()));
  PICARD_TRY(up(e->A.inflow_map, imap.data(), sizeof(Tile) * imap.size()));
  PICARD_TRY(up(e->A.g, &gl, sizeof(Global)));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  e->dirty = false;
  return PICARD_OK;
}

// ---- launch sequences (captured into graphs, or launched directly)

static int grid_particles(picard_ensemble *e) { return 8 * e->cus; }

static int launch_pre(picard_ensemble *e, int cur) {
  hipStream_t st = e->stream;
  const Arenas &A = e->A;
  const PartArena &X = e->P[cur], &Y = e->P[cur ^ 1];
  const int gp = grid_particles(e);
  k_inflow<<<gp, kT, 0, st>>>(A, X, e->cap);
  k_inflow_commit<<<1, 64, 0, st>>>(A, e->cap);
  k_push<<<gp, kT, 0, st>>>(A, X);
  k_fill_i32<<<1024, kT, 0, st>>>(A.hist, A.kcap + 1, 0);
  k_fill_i32<<<1024, kT, 0, st>>>(A.cursor, A.kcap + 1, 0);
  k_flag<<<gp, kT, 0, st>>>(A, X);
  PICARD_HIP(hipGetLastError());
  PICARD_TRY(scan_exclusive_stream(st, A.hist, A.hist, A.kcap, e->scan_tmp, e->scan_tmp_len));
  k_scatter<<<(int)((e->cap + kT - 1) / kT), kT, 0, st>>>(A, X, Y);
  k_ranges<<<(kMaxSlots + kT - 1) / kT, kT, 0, st>>>(A);
  k_fill_u64<<<1024, kT, 0, st>>>(A.Nfix, 2 * e->max_nodes);
  k_deposit<<<gp, kT, 0, st>>>(A, Y);
  k_fold<<<4 * e->cus, kT, 0, st>>>(A);
  k_bnorm<<<e->cus, kT, 0, st>>>(A);
  k_facesum<<<4 * e->cus, kT, 0, st>>>(A);
  k_facereduce<<<4 * e->cus, kT, 0, st>>>(A);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

// Sweep jj of the chunk; after every kDecideEvery-th, the stop decision over
// the group's iterates.
template <int CPT>
static void launch_sweep(picard_ensemble *e, bool tune, int jj) {
  const int blocks = 8 * ((e->tile.blocks_per_cu * e->cus + 7) / 8);
  if (tune) {
    k_sweep<CPT, true><<<blocks, kT, 0, e->stream>>>(e->A, e->inner, jj);
  } else {
    k_sweep<CPT, false><<<blocks, kT, 0, e->stream>>>(e->A, e->inner, jj);
    if ((jj + 1) % kDecideEvery == 0)
      k_decide<<<e->cus, kT, 0, e->stream>>>(e->A, e->inner, jj + 1 - kDecideEvery);
  }
}

static int launch_sweeps(picard_ensemble *e, int count, bool tune) {
  for (int q = 0; q < count; ++q) {
    switch (e->tile.cells_per_thread) {
      case 1: launch_sweep<1>(e, tune, q); break;
      case 2: launch_sweep<2>(e, tune, q); break;
      case 4: launch_sweep<4>(e, tune, q); break;
      default: launch_sweep<8>(e, tune, q); break;
    }
  }
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

static int launch_chunk(picard_ensemble *e, bool with_outer) {
  const int blocks = 8 * ((e->tile.blocks_per_cu * e->cus + 7) / 8);
  if (with_outer) {
    k_outer<<<blocks, kT, 0, e->stream>>>(e->A, e->tile.cells_per_thread);
    k_outer_done<<<(kMaxSlots + kT - 1) / kT, kT, 0, e->stream>>>(e->A);
  }
  PICARD_TRY(launch_sweeps(e, kChunkSweeps, false));
  k_count_active<<<1, kT, 0, e->stream>>>(e->A, e->host_active);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

static int launch_post(picard_ensemble *e, int cur) {
  hipStream_t st = e->stream;
  const Arenas &A = e->A;
  k_finish<<<4 * e->cus, kT, 0, st>>>(A);
  k_efield<<<4 * e->cus, kT, 0, st>>>(A);
  k_boris<<<grid_particles(e), kT, 0, st>>>(A, e->P[cur ^ 1]);
  k_step_end<<<1, kT, 0, st>>>(A, e->host_status, e->host_n, e->host_err);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

template <typename F>
static int capture(picard_ensemble *e, hipGraphExec_t *out, F &&f) {
  hipGraph_t g = nullptr;
  PICARD_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  const int rc = f();
  hipGraph_t got = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->stream, &got);
  g = got;
  if (rc < 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) {
    set_error(std::string("ensemble: graph capture failed: ") + hipGetErrorString(ec));
    return PICARD_ERR_HIP;
  }
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    set_error(std::string("ensemble: graph instantiation failed: ") + hipGetErrorString(ei));
    return PICARD_ERR_HIP;
  }
  return PICARD_OK;
}

static int ensure_graphs(picard_ensemble *e) {
  if (!e->cfg.use_graphs || e->g_pre[0]) return PICARD_OK;
  for (int p = 0; p < 2; ++p) {
    PICARD_TRY(capture(e, &e->g_pre[p], [&] { return launch_pre(e, p); }));
    PICARD_TRY(capture(e, &e->g_post[p], [&] { return launch_post(e, p); }));
  }
  PICARD_TRY(capture(e, &e->g_chunk[0], [&] { return launch_chunk(e, false); }));
  PICARD_TRY(capture(e, &e->g_chunk[1], [&] { return launch_chunk(e, true); }));
  return PICARD_OK;
}

#define ENS_CHECK(e)                              \
  do {                                            \
    if (!(e)) {                                   \
      set_error("ensemble: NULL handle");         \
      return PICARD_ERR_ARG;                      \
    }                                             \
    PICARD_HIP(hipSetDevice((e)->device));        \
  } while (0)

extern "C" {

int picard_sim_derive(const picard_sim_params *p, picard_sim_derived *out) {
  return derive_sim(p, out);
}

int picard_sim_footprint(const picard_sim_params *p, int64_t *grid_bytes, int64_t *particles) {
  picard_sim_derived d;
  PICARD_TRY(derive_sim(p, &d));
  int64_t nodes, cells, dist, face, keys;
  int32_t fch;
  slot_sizes(p, &nodes, &cells, &dist, &face, &keys, &fch);
  // U, u0, u1, C (FP64), N, Nfix (2 species), E32 (16 B), r32, the
  // correction buffers, distance table, face partials, keys (hist + cursor)
  const int64_t b = nodes * (8 * 4 + 16 + 16 + 16) + cells * (4 + 4 * kEBufs) + dist * 8 +
                    face * 8 + keys * 8;
  if (grid_bytes) *grid_bytes = b;
  if (particles)
    *particles = cells * 2 * p->ppc + 2 * (int64_t)p->ny * p->nz * 2 * p->ppc;
  return PICARD_OK;
}

int picard_ensemble_create(const picard_ensemble_config *cfg, picard_ensemble **out) {
  if (!cfg || !out) {
    set_error("picard_ensemble_create: NULL");
    return PICARD_ERR_ARG;
  }
  *out = nullptr;
  if (cfg->max_sims < 1 || cfg->max_sims > kMaxSlots || cfg->max_cells < 8 ||
      cfg->max_particles < 1 || cfg->max_particles > ((int64_t)1 << 31)) {
    set_error("picard_ensemble_create: max_sims in 1..4096, max_cells >= 8, max_particles in "
              "1..2^31");
    return PICARD_ERR_ARG;
  }
  const int inner = cfg->inner == 0 ? 512 : cfg->inner;
  if (inner < kChunkSweeps || inner > kMaxSplitInner || inner % kChunkSweeps) {
    set_error("picard_ensemble_create: inner must be a multiple of 64 in 64..512");
    return PICARD_ERR_ARG;
  }
  struct Guard {
    picard_ensemble *p;
    ~Guard() {
      if (p) (void)picard_ensemble_destroy(p);
    }
  };
  picard_ensemble *e = new picard_ensemble();
  Guard guard{e};
  e->cfg = *cfg;
  e->inner = inner;
  e->device = cfg->device;
  PICARD_HIP(hipSetDevice(e->device));
  hipDeviceProp_t prop;
  PICARD_HIP(hipGetDeviceProperties(&prop, e->device));
  e->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  PICARD_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  e->cap = cfg->max_particles;
  e->max_cells = cfg->max_cells;
  // node, distance, face and key capacities follow from the cell capacity:
  // with at least 8 cells per axis a slot has at most (10/8)^3 < 2 nodes,
  // (9/8)^3 < 1.5 distance entries and 1/2 face partials per cell (x-planes in
  // chunks of 8), and twice its distance entries as keys
  e->max_nodes = 2 * cfg->max_cells + 1024;
  e->max_dist = 3 * (cfg->max_cells / 2) + 1024;
  e->max_face = cfg->max_cells / 2 + 1024 * (int64_t)cfg->max_sims;
  const int64_t kcap = 2 * e->max_dist;
  if (kcap > INT32_MAX - 1) {
    set_error("picard_ensemble_create: max_cells too large for 32-bit sort keys");
    return PICARD_ERR_ARG;
  }
  e->A.kcap = kcap;
  e->slots.resize(cfg->max_sims);
  e->ag.reset(e->max_nodes);
  e->ac.reset(e->max_cells);
  e->ad.reset(e->max_dist);
  e->af.reset(e->max_face);
  e->ak.reset(kcap);
  Arenas &A = e->A;
  const int S = cfg->max_sims;
  PICARD_TRY(dalloc_n(e, &A.sims, S));
  PICARD_TRY(dalloc_n(e, &A.state, S));
  PICARD_TRY(dalloc_n(e, &A.g, 1));
  PICARD_TRY(dalloc_n(e, &A.live, S));
  PICARD_TRY(dalloc_n(e, &A.inflow_map, S));
  // tile lists sized for the capacities (smallest tiles: 64 threads x 1 cell)
  e->cap_node_t = e->max_nodes / kT + S;
  e->cap_cell_t = e->max_cells / 64 + 8 * S;
  e->cap_face_t = 2 * (e->max_face / kT) + 1024 * (int64_t)S;
  e->cap_fred_t = e->max_face / kT + 2 * S;
  PICARD_TRY(dalloc_n(e, &A.node_tiles, e->cap_node_t));
  PICARD_TRY(dalloc_n(e, &A.cell_tiles, e->cap_cell_t));
  PICARD_TRY(dalloc_n(e, &A.units, kDecideEvery * e->cap_cell_t));
  A.ucap = e->cap_cell_t;
  PICARD_TRY(dalloc_n(e, &A.face_tiles, e->cap_face_t));
  PICARD_TRY(dalloc_n(e, &A.fred_tiles, e->cap_fred_t));
  PICARD_TRY(dalloc_n(e, &A.U, e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.u[0], e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.u[1], e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.C, e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.N, 2 * e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.Nfix, 2 * e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.E32, e->max_nodes));
  PICARD_TRY(dalloc_n(e, &A.D, e->max_dist));
  PICARD_TRY(dalloc_n(e, &A.fscr, e->max_face));
  PICARD_TRY(dalloc_n(e, &A.r32, e->max_cells));
  for (int q = 0; q < kEBufs; ++q) PICARD_TRY(dalloc_n(e, &A.e32[q], e->max_cells));
  PICARD_TRY(dalloc_n(e, &A.hist, kcap + 1));
  PICARD_TRY(dalloc_n(e, &A.cursor, kcap + 1));
  PICARD_TRY(dalloc_n(e, &A.key, e->cap));
  e->scan_tmp_len = scan_tmp_required(kcap);
  PICARD_TRY(dalloc_n(e, &e->scan_tmp, e->scan_tmp_len));
  for (int q = 0; q < 2; ++q) {
    PartArena &p = e->P[q];
    for (int d = 0; d < 3; ++d) {
      PICARD_TRY(dalloc_n(e, &p.c[d], e->cap));
      PICARD_TRY(dalloc_n(e, &p.f[d], e->cap));
      PICARD_TRY(dalloc_n(e, &p.v[d], e->cap));
    }
    PICARD_TRY(dalloc_n(e, &p.sp, e->cap));
    PICARD_TRY(dalloc_n(e, &p.sim, e->cap));
    PICARD_TRY(dalloc_n(e, &p.id, e->cap));
  }
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->host_active), sizeof(int32_t),
                           hipHostMallocMapped | hipHostMallocCoherent));
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->host_status),
                           sizeof(picard_sim_status) * S, hipHostMallocMapped | hipHostMallocCoherent));
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->host_n), sizeof(int64_t),
                           hipHostMallocMapped | hipHostMallocCoherent));
  PICARD_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->host_err), sizeof(int32_t),
                           hipHostMallocMapped | hipHostMallocCoherent));
  for (auto &ev : e->chunk_ev) PICARD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  std::memset(e->host_status, 0, sizeof(picard_sim_status) * S);
  *e->host_n = 0;
  *e->host_err = 0;
  *e->host_active = 0;
  // device tables start zeroed (fill kernels, no memset nodes)
  std::vector<SimState> st0(S);
  std::memset(st0.data(), 0, sizeof(SimState) * S);
  PICARD_HIP(hipMemcpyAsync(A.state, st0.data(), sizeof(SimState) * S, hipMemcpyHostToDevice,
                            e->stream));
  Global g0{};
  PICARD_HIP(hipMemcpyAsync(A.g, &g0, sizeof(Global), hipMemcpyHostToDevice, e->stream));
  k_fill_f32<<<1024, kT, 0, e->stream>>>(A.r32, e->max_cells);
  for (int q = 0; q < kEBufs; ++q) k_fill_f32<<<1024, kT, 0, e->stream>>>(A.e32[q], e->max_cells);
  k_fill_u64<<<1024, kT, 0, e->stream>>>(reinterpret_cast<unsigned long long *>(A.U),
                                         e->max_nodes);
  k_fill_u64<<<1024, kT, 0, e->stream>>>(reinterpret_cast<unsigned long long *>(A.E32),
                                         2 * e->max_nodes);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipStreamSynchronize(e->stream));
  PICARD_TRY(rebuild(e));
  guard.p = nullptr;
  *out = e;
  return PICARD_OK;
}

int picard_ensemble_destroy(picard_ensemble *e) {
  if (!e) return PICARD_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_graphs(e);
  for (void *p : e->allocs) (void)hipFree(p);
  for (void *p : {(void *)e->host_active, (void *)e->host_status, (void *)e->host_n,
                  (void *)e->host_err})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t ev : e->chunk_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PICARD_OK;
}

int picard_ensemble_add(picard_ensemble *e, const picard_sim_params *p, int32_t *slot) {
  ENS_CHECK(e);
  picard_sim_derived d;
  PICARD_TRY(derive_sim(p, &d));
  if (p->nx > 32000 || p->ny > 32000 || p->nz > 32000) {
    set_error("picard_ensemble_add: int16 cell coordinates");
    return PICARD_ERR_ARG;
  }
  int s = -1;
  for (int q = 0; q < (int)e->slots.size(); ++q)
    if (!e->slots[q].live && !e->slots[q].draining) {
      s = q;
      break;
    }
  if (s < 0) {
    set_error("picard_ensemble_add: no free slot");
    return PICARD_ERR_CAPACITY;
  }
  int64_t nodes, cells, dist, face, keys;
  int32_t fch;
  slot_sizes(p, &nodes, &cells, &dist, &face, &keys, &fch);
  const int64_t init = cells * 2 * p->ppc;
  const int64_t inflow = 2 * (int64_t)p->ny * p->nz * 2 * p->ppc;
  if (e->n_host + init + inflow > e->cap) {
    set_error("picard_ensemble_add: particle capacity exceeded");
    return PICARD_ERR_CAPACITY;
  }
  SlotHost h;
  h.p = *p;
  h.d = d;
  h.g.len = nodes;
  h.c.len = cells;
  h.dt.len = dist;
  h.f.len = face;
  h.k.len = keys;
  // reserve the slot's ranges; on a failure give back what was taken
  struct Want {
    Allocator *a;
    Range *r;
  } wants[5] = {{&e->ag, &h.g}, {&e->ac, &h.c}, {&e->ad, &h.dt}, {&e->af, &h.f}, {&e->ak, &h.k}};
  int got = 0;
  for (; got < 5; ++got)
    if (!wants[got].a->alloc(wants[got].r->len, &wants[got].r->off)) break;
  if (got < 5) {
    for (int q = 0; q < got; ++q) wants[q].a->release(wants[q].r->off, wants[q].r->len);
    set_error("picard_ensemble_add: grid capacity exceeded");
    return PICARD_ERR_CAPACITY;
  }
  h.live = true;
  h.init_particles = init;
  e->slots[s] = h;
  const int rb = rebuild(e);
  if (rb < 0) {                      // work lists full: give the slot back
    for (auto &w : wants) w.a->release(w.r->off, w.r->len);
    e->slots[s] = SlotHost();
    (void)rebuild(e);
    return rb;
  }
  // fresh state: step 0, ids from 0
  SimState st0;
  std::memset(&st0, 0, sizeof(st0));
  st0.next_id = init;
  st0.start = e->n_host;
  st0.count = init;
  PICARD_HIP(hipMemcpyAsync(&e->A.state[s], &st0, sizeof(SimState), hipMemcpyHostToDevice,
                            e->stream));
  k_distance<<<e->cus, kT, 0, e->stream>>>(e->A, s);
  k_initial<<<4 * e->cus, kT, 0, e->stream>>>(e->A, e->P[e->cur], s, e->n_host);
  PICARD_HIP(hipGetLastError());
  e->n_host += init;
  Global gl;
  PICARD_HIP(hipMemcpyAsync(&gl, e->A.g, sizeof(Global), hipMemcpyDeviceToHost, e->stream));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  gl.n = e->n_host;
  PICARD_HIP(hipMemcpyAsync(e->A.g, &gl, sizeof(Global), hipMemcpyHostToDevice, e->stream));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  e->slots[s].appended_at = e->n_host - init;
  picard_sim_status o{};
  o.live = 1;
  o.num_local = init;
  e->host_status[s] = o;
  if (slot) *slot = s;
  return PICARD_OK;
}

int picard_ensemble_remove(picard_ensemble *e, int32_t slot) {
  ENS_CHECK(e);
  if (slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].live) {
    set_error("picard_ensemble_remove: not a live slot");
    return PICARD_ERR_ARG;
  }
  SlotHost &h = e->slots[slot];
  e->ag.release(h.g.off, h.g.len);
  e->ac.release(h.c.off, h.c.len);
  e->ad.release(h.dt.off, h.dt.len);
  e->af.release(h.f.off, h.f.len);
  e->ak.release(h.k.off, h.k.len);
  h = SlotHost();
  h.draining = true;
  e->host_status[slot] = picard_sim_status{};
  PICARD_TRY(rebuild(e));   // the slot's SimDev now has live = 0: the next sort drops its particles
  return PICARD_OK;
}

int picard_ensemble_step(picard_ensemble *e, picard_ensemble_stats *stats) {
  ENS_CHECK(e);
  if (e->dirty) PICARD_TRY(rebuild(e));
  PICARD_TRY(ensure_graphs(e));
  int live = 0;
  for (auto &s : e->slots) live += s.live ? 1 : 0;
  picard_ensemble_stats out{};
  if (live == 0) {
    if (stats) *stats = out;
    return PICARD_OK;
  }
  const int cur = e->cur;
  const auto t_start = std::chrono::steady_clock::now();
  if (e->cfg.use_graphs) PICARD_HIP(hipGraphLaunch(e->g_pre[cur], e->stream));
  else PICARD_TRY(launch_pre(e, cur));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  const auto t_solve = std::chrono::steady_clock::now();
  // the solve: replays of 64-sweep chunks, a refinement pass every `inner`
  // sweeps, until no simulation is solving
  const int per_ref = e->inner / kChunkSweeps;
  // every solve ends within 2*maxit+1 sweeps; the bound only guards the
  // host loop against a device state that never reports the end
  int64_t maxit = 1;
  for (auto &s : e->slots)
    if (s.live) maxit = std::max<int64_t>(maxit, s.p.maxit);
  const int64_t max_chunks = (2 * maxit + 2) / kChunkSweeps + 2 * per_ref + 2;
  // One chunk is kept enqueued beyond the one the host waits for, so the
  // device never idles on the host's read of the count; the chunk enqueued
  // after the last one needed finds no simulation solving and changes no
  // state (its sweeps and decisions skip stopped simulations, and a
  // refinement pass touches solving ones only).
  int64_t chunks = 0, enqueued = 0;
  auto enqueue = [&]() -> int {
    const bool outer = enqueued % per_ref == 0;
    if (e->cfg.use_graphs) PICARD_HIP(hipGraphLaunch(e->g_chunk[outer ? 1 : 0], e->stream));
    else PICARD_TRY(launch_chunk(e, outer));
    PICARD_HIP(hipEventRecord(e->chunk_ev[enqueued % 2], e->stream));
    ++enqueued;
    return PICARD_OK;
  };
  PICARD_TRY(enqueue());
  for (;;) {
    if (enqueued < max_chunks) PICARD_TRY(enqueue());
    PICARD_HIP(hipEventSynchronize(e->chunk_ev[chunks % 2]));
    ++chunks;
    if (*reinterpret_cast<volatile int32_t *>(e->host_active) == 0) break;
    if (chunks >= max_chunks) {
      set_error("ensemble step: the solve did not report its end within 2*maxit+1 sweeps");
      (void)hipStreamSynchronize(e->stream);
      return PICARD_ERR_STATE;
    }
  }
  const auto t_post = std::chrono::steady_clock::now();
  if (e->cfg.use_graphs) PICARD_HIP(hipGraphLaunch(e->g_post[cur], e->stream));
  else PICARD_TRY(launch_post(e, cur));
  PICARD_HIP(hipStreamSynchronize(e->stream));
  const auto t_end = std::chrono::steady_clock::now();
  out.step_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
  out.solve_ms = std::chrono::duration<double, std::milli>(t_post - t_solve).count();
  e->cur ^= 1;
  e->n_host = *e->host_n;
  for (auto &s : e->slots) {
    s.appended_at = -1;
    s.draining = false;
  }
  out.sims = live;
  out.particles = 0;
  out.solve_chunks = enqueued;
  for (int s = 0; s < (int)e->slots.size(); ++s) {
    if (!e->slots[s].live) continue;
    const picard_sim_status &o = e->host_status[s];
    out.noconv += o.noconv;
    out.particles += o.num_local;
    out.pushed += o.pushed;
    out.max_sweeps = std::max<int64_t>(out.max_sweeps, o.sweeps);
    out.sum_sweeps += o.sweeps;
    out.cell_sweeps += o.sweeps * (int64_t)e->slots[s].c.len;
  }
  if (stats) *stats = out;
  const int32_t err = *e->host_err;
  if (err & 1) {
    set_error("ensemble step: particle capacity exceeded (inflow not appended)");
    return PICARD_ERR_CAPACITY;
  }
  if (err & 2) {
    set_error("ensemble step: particles outside their grid were dropped");
    return PICARD_ERR_RANGE;
  }
  return out.noconv ? PICARD_WARN_NOCONV : PICARD_OK;
}

int picard_ensemble_status(picard_ensemble *e, int32_t slot, picard_sim_status *out) {
  ENS_CHECK(e);
  if (slot < 0 || slot >= (int)e->slots.size() || !out) {
    set_error("picard_ensemble_status: bad slot");
    return PICARD_ERR_ARG;
  }
  *out = e->host_status[slot];
  return PICARD_OK;
}

int picard_ensemble_field_read(picard_ensemble *e, int32_t slot, int32_t which, double *host,
                               int64_t count) {
  ENS_CHECK(e);
  if (slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].live || !host) {
    set_error("picard_ensemble_field_read: not a live slot");
    return PICARD_ERR_ARG;
  }
  const SlotHost &h = e->slots[slot];
  const int64_t G = h.g.len;
  const int64_t want = which == PICARD_FIELD_E ? 3 * G : which == PICARD_FIELD_N ? 2 * G : G;
  if ((which != PICARD_FIELD_U && which != PICARD_FIELD_E && which != PICARD_FIELD_N &&
       which != PICARD_FIELD_C) || count != want) {
    set_error("picard_ensemble_field_read: field U, E, N or C with its element count");
    return PICARD_ERR_ARG;
  }
  const double *src = nullptr;
  double *tmp = nullptr;
  if (which == PICARD_FIELD_U) src = e->A.U + h.g.off;
  else if (which == PICARD_FIELD_C) src = e->A.C + h.g.off;
  else if (which == PICARD_FIELD_N) src = e->A.N + 2 * h.g.off;
  else {
    PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&tmp), sizeof(double) * 3 * G));
    k_widen<<<256, kT, 0, e->stream>>>(e->A.E32 + h.g.off, G, tmp);
    src = tmp;
  }
  const hipError_t ec = hipMemcpyAsync(host, src, sizeof(double) * want, hipMemcpyDeviceToHost,
                                       e->stream);
  const hipError_t es = hipStreamSynchronize(e->stream);
  if (tmp) (void)hipFree(tmp);
  if (ec != hipSuccess || es != hipSuccess) {
    set_error("picard_ensemble_field_read: copy failed");
    return PICARD_ERR_HIP;
  }
  return PICARD_OK;
}

int picard_ensemble_particles_read(picard_ensemble *e, int32_t slot, double *rp, int32_t *sp,
                                   int64_t count) {
  ENS_CHECK(e);
  if (slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].live || !rp || !sp) {
    set_error("picard_ensemble_particles_read: not a live slot");
    return PICARD_ERR_ARG;
  }
  const SlotHost &h = e->slots[slot];
  int64_t first, n;
  if (h.appended_at >= 0) {
    first = h.appended_at;
    n = h.init_particles;
  } else {
    SimState st;
    PICARD_HIP(hipMemcpyAsync(&st, &e->A.state[slot], sizeof(SimState), hipMemcpyDeviceToHost,
                              e->stream));
    PICARD_HIP(hipStreamSynchronize(e->stream));
    first = st.start;
    n = st.count;
  }
  if (count != n) {
    set_error("picard_ensemble_particles_read: count must equal the simulation's particles (" +
              std::to_string(n) + ")");
    return PICARD_ERR_ARG;
  }
  if (n == 0) return PICARD_OK;
  double *drp = nullptr;
  int32_t *dsp = nullptr;
  PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&drp), sizeof(double) * 8 * n));
  if (hipMalloc(reinterpret_cast<void **>(&dsp), sizeof(int32_t) * n) != hipSuccess) {
    (void)hipFree(drp);
    set_error("picard_ensemble_particles_read: out of device memory");
    return PICARD_ERR_HIP;
  }
  k_to_aos<<<256, kT, 0, e->stream>>>(e->A.sims, e->P[e->cur], slot, first, n, drp, dsp);
  hipError_t ec = hipMemcpyAsync(rp, drp, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, e->stream);
  if (ec == hipSuccess)
    ec = hipMemcpyAsync(sp, dsp, sizeof(int32_t) * n, hipMemcpyDeviceToHost, e->stream);
  const hipError_t es = hipStreamSynchronize(e->stream);
  (void)hipFree(drp);
  (void)hipFree(dsp);
  if (ec != hipSuccess || es != hipSuccess) {
    set_error("picard_ensemble_particles_read: copy failed");
    return PICARD_ERR_HIP;
  }
  return PICARD_OK;
}

int picard_ensemble_set_tile(picard_ensemble *e, const picard_ensemble_tile *t) {
  ENS_CHECK(e);
  if (!t || (t->cells_per_thread != 1 && t->cells_per_thread != 2 && t->cells_per_thread != 4 &&
             t->cells_per_thread != 8) ||
      t->threads != kT || t->blocks_per_cu < 1 || t->blocks_per_cu > 8) {
    set_error("picard_ensemble_set_tile: cells_per_thread 1, 2, 4 or 8; threads 256; "
              "blocks_per_cu 1..8");
    return PICARD_ERR_ARG;
  }
  e->tile = *t;
  e->tile.reserved = 0;
  free_graphs(e);
  return rebuild(e);
}

int picard_ensemble_get_tile(picard_ensemble *e, picard_ensemble_tile *t) {
  ENS_CHECK(e);
  if (!t) {
    set_error("picard_ensemble_get_tile: NULL");
    return PICARD_ERR_ARG;
  }
  *t = e->tile;
  return PICARD_OK;
}

int picard_ensemble_tune(picard_ensemble *e, int32_t sweeps, picard_ensemble_tile *chosen,
                         double *times_us, int32_t max_times, int32_t *ncandidates) {
  ENS_CHECK(e);
  if (sweeps < 1 || sweeps > 100000) {
    set_error("picard_ensemble_tune: sweeps in 1..100000");
    return PICARD_ERR_ARG;
  }
  if (e->dirty) PICARD_TRY(rebuild(e));
  std::vector<picard_ensemble_tile> cand;
  for (int cpt : {1, 2, 4, 8})
    for (int bpc : {2, 4, 8}) cand.push_back({cpt, kT, bpc, 0});
  const picard_ensemble_tile keep = e->tile;
  hipEvent_t t0, t1;
  PICARD_HIP(hipEventCreate(&t0));
  PICARD_HIP(hipEventCreate(&t1));
  int best = -1;
  double best_us = 0.0;
  int rc = PICARD_OK;
  for (size_t q = 0; q < cand.size(); ++q) {
    e->tile = cand[q];
    rc = rebuild(e);
    if (rc < 0) break;
    // warm-up launch, then the timed ones (tuning launches skip the stop test
    // and every solve state; they only rewrite the scratch corrections)
    rc = launch_sweeps(e, 1, true);
    if (rc < 0) break;
    if (hipEventRecord(t0, e->stream) != hipSuccess) { rc = PICARD_ERR_HIP; break; }
    rc = launch_sweeps(e, sweeps, true);
    if (rc < 0) break;
    if (hipEventRecord(t1, e->stream) != hipSuccess ||
        hipEventSynchronize(t1) != hipSuccess) { rc = PICARD_ERR_HIP; break; }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    const double us = 1e3 * ms / sweeps;
    if (times_us && (int32_t)q < max_times) times_us[q] = us;
    if (best < 0 || us < best_us) {
      best = (int)q;
      best_us = us;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc < 0) {
    e->tile = keep;
    (void)rebuild(e);
    if (rc == PICARD_ERR_HIP) set_error("picard_ensemble_tune: timing failed");
    return rc;
  }
  if (ncandidates) *ncandidates = (int32_t)cand.size();
  e->tile = cand[best];
  free_graphs(e);
  PICARD_TRY(rebuild(e));
  if (chosen) *chosen = e->tile;
  return PICARD_OK;
}

}  // extern "C"
