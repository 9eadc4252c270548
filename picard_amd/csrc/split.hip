// split.hip -- the FP32-split form of the time step.
//
// Particles: int16 lower-node coordinates (i, j, k, 0-based) plus FP32
// in-cell offsets f = xi - floor(xi), xi = (x - Lmin)/dx + 1/2 (so the CIC
// weights of ParticlesGridify.f90:70-88 are f and 1 - f), FP32 velocities,
// a uint8 species and a uint32 id.  Every step re-sorts the particles by cell
// (counting sort), which also drops the ones BCoutflow deletes, and replaces
// ParticlesTransfer's single-domain wrap by index arithmetic.
//
// Deposit: int64 fixed point (2^-40 units) with integer guard folds, so the
// densities do not depend on particle order.  Field solve: iterative
// refinement -- FP32 Jacobi sweeps on the correction e of A e = r with an
// FP64 outer residual r = b - A u and FP64 update u += e; the residual norm of
// every sweep's iterate u + e_j (= r - A e_j) is accumulated in FP64, so the
// stop rule of CalcEpotential.f90:60-62 is applied sweep by sweep.  E is
// produced in FP32 for the gather; BCpotentials runs in FP64 (grid.hip).
#include <hip/hip_ext.h>

#include <cmath>

#include "cic.h"
#include "picard_internal.h"
#include "reduce.h"
#include "sched.h"

namespace picard {
namespace {

constexpr int kT = 256;
// threads per block of the persistent stencil kernels (4 waves: a 64 x 4
// tile; eight blocks per CU at 8 waves per SIMD; 1024-thread blocks measured
// slower)
constexpr int kStencilT = 256;
constexpr double kFix = 1099511627776.0;       // 2^40
constexpr double kFixInv = 1.0 / 1099511627776.0;

__device__ __forceinline__ int64_t lin_key(const GridDims &g, int i, int j, int k) {
  return (int64_t)i + (int64_t)(g.nx + 1) * ((int64_t)j + (int64_t)(g.ny + 1) * k);
}

// Carry a coordinate offset into [0, 1): returns the integer carry.
__device__ __forceinline__ int carry(float &f) {
  const float fl = floorf(f);
  f = f - fl;
  int c = (int)fl;
  if (f >= 1.0f) {           // f - floor(f) rounded up to 1
    f = 0.0f;
    c += 1;
  }
  return c;
}

// AdvancePosition in the split form: xi += v dt/dx.
__global__ void __launch_bounds__(kT) k_sp_push(ParticlesSplit p, int64_t n, float3 dtdx) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= n) return;
  float fx = p.f[0][q] + p.v[0][q] * dtdx.x;
  float fy = p.f[1][q] + p.v[1][q] * dtdx.y;
  float fz = p.f[2][q] + p.v[2][q] * dtdx.z;
  p.c[0][q] = (int16_t)(p.c[0][q] + carry(fx));
  p.c[1][q] = (int16_t)(p.c[1][q] + carry(fy));
  p.c[2][q] = (int16_t)(p.c[2][q] + carry(fz));
  p.f[0][q] = fx;
  p.f[1][q] = fy;
  p.f[2][q] = fz;
}

// xi < 1/2 is x < Lmin; xi > n + 1/2 is x > Lmax.
__device__ __forceinline__ bool below(int c, float f) { return c < 0 || (c == 0 && f < 0.5f); }
__device__ __forceinline__ bool above(int c, float f, int n) {
  return c > n || (c == n && f > 0.5f);
}

// BCoutflow.f90:28-66 on the split coordinates: flow faces, and the band one
// domain width beyond the box in every dimension.
__device__ __forceinline__ bool sp_delete(const GridDims &g, const int c[3], const float f[3]) {
  const int n[3] = {g.nx, g.ny, g.nz};
  bool del = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (g.flow[d] && (below(c[d], f[d]) || above(c[d], f[d], n[d]))) del = true;
    if (below(c[d] + n[d], f[d]) || above(c[d] - n[d], f[d], n[d])) del = true;
  }
  return del;
}

// ParticlesTransfer.f90:494-519: periodic wrap by +-n cells.
__device__ __forceinline__ void sp_wrap(const GridDims &g, int c[3], const float f[3]) {
  const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int d = 2; d >= 0; --d) {
    if (below(c[d], f[d])) c[d] += n[d];
    if (above(c[d], f[d], n[d])) c[d] -= n[d];
  }
}

// counter[k] += 1 for every lane of the wave with k >= 0, one atomic per
// distinct key: the lanes sharing the leader's key (the lowest pending lane)
// take ranks by lane order after the value the leader's add of their count
// returned.  Returns what the lane's own add of 1 would have returned in one
// order of the adds (-1 for k < 0).  Particles arrive nearly sorted by key,
// so a wave holds two or three keys and the adds do not meet on one address
// lane after lane.  Every lane stays in the loop (its condition is the
// wave-uniform mask of pending lanes) and values cross lanes only through
// readlane, which reads the lane whatever the exec mask; the leader's bit
// leaves the mask every pass, so the loop ends within 64 passes.
__device__ __forceinline__ int32_t wave_add1(int32_t *counter, int32_t k) {
  const int lane = (int)(threadIdx.x & 63);
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

// Sort key of a kept particle (its lower node and species), -1 for a deleted one.
__global__ void __launch_bounds__(kT) k_sp_flag(ParticlesSplit p, int64_t n, GridDims g,
                                                 int wrap, int32_t *__restrict__ key,
                                                 int32_t *__restrict__ hist) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  int32_t k = -1;                          // lanes past n stay for the wave's adds
  if (q < n) {
    int c[3] = {p.c[0][q], p.c[1][q], p.c[2][q]};
    const float f[3] = {p.f[0][q], p.f[1][q], p.f[2][q]};
    if (!sp_delete(g, c, f)) {
      if (wrap) {
        sp_wrap(g, c, f);
        p.c[0][q] = (int16_t)c[0];
        p.c[1][q] = (int16_t)c[1];
        p.c[2][q] = (int16_t)c[2];
      }
      const bool in = c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] <= g.nx && c[1] <= g.ny &&
                      c[2] <= g.nz;
      // key (cell, species): the deposit then sees runs of one cell and species
      const int32_t cells = (int32_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
      k = in ? (int32_t)lin_key(g, c[0], c[1], c[2]) * g.S + ((int32_t)p.sp[q] - 1) : cells * g.S;
    }
    key[q] = k;
  }
  wave_add1(hist, k);
}

__global__ void __launch_bounds__(kT) k_sp_scatter(ParticlesSplit src, ParticlesSplit dst, int64_t n,
                                                    const int32_t *__restrict__ key,
                                                    const int32_t *__restrict__ off,
                                                    int32_t *__restrict__ cursor) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int32_t k = q < n ? key[q] : -1;   // lanes past n stay for the wave's adds
  const int32_t r = wave_add1(cursor, k);
  if (k < 0 || r < 0) return;
  const int64_t at = (int64_t)off[k] + r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dst.c[d][at] = src.c[d][q];
    dst.f[d][at] = src.f[d][q];
    dst.v[d][at] = src.v[d][q];
  }
  dst.sp[at] = src.sp[q];
  dst.id[at] = src.id[q];
}

// Fixed-point CIC deposit (int64 atomics, order independent).  FLUX adds
// the number fluxes of ParticlesGridifyAll (ParticlesGridifyAll.f90:126-212)
// as weight * v_c in units of 2^-24 m/s (Ffix: s fastest, then c, then node).
constexpr double kFluxFix = 16777216.0;         // 2^24
constexpr double kFluxFixInv = 1.0 / 16777216.0;

// Particles are sorted by (cell, species), so a thread walks kDepChunk
// consecutive particles and adds the fixed-point weights of each run of one
// cell and species in registers, flushing one add per corner when the run
// ends: integer sums, so the result equals one atomic per particle.  Without
// the fluxes the adds are staged in LDS per block: the block's 4096 sorted
// particles cover a short stretch of lower nodes from the first one's, b0, so
// their corners fall in four node segments b0 + {0, X, XY, XY + X} + [0,
// kDepSeg); runs whose corners lie inside go to the segments with LDS
// atomics, the rest straight to memory, and the block then adds each nonzero
// segment entry to memory once (a few hundred global atomics per block
// instead of eight per run).
constexpr int kDepChunk = 16;
constexpr int kDepSeg = 128;                    // nodes per LDS segment
constexpr int kDepLdsSpecies = 3;               // species the LDS segments hold

template <bool FLUX>
__global__ void __launch_bounds__(kT) k_sp_deposit(ParticlesSplit p, int64_t n, GridDims g,
                                                    unsigned long long *__restrict__ Nfix,
                                                    unsigned long long *__restrict__ Ffix,
                                                    int32_t *__restrict__ errors) {
  __shared__ unsigned long long seg[FLUX ? 1 : 4 * kDepSeg * kDepLdsSpecies];
  __shared__ int64_t seg_b0;
  __shared__ int seg_on;
  const int64_t q0 = ((int64_t)blockIdx.x * kT + threadIdx.x) * kDepChunk;
  if (FLUX && q0 >= n) return;
  const int64_t q1 = q0 + kDepChunk < n ? q0 + kDepChunk : n;   // empty past n
  const int64_t X = g.X, XY = (int64_t)g.X * g.Y;
  const int S = g.S;
  int64_t b0 = 0;
  bool on = false;
  if constexpr (!FLUX) {
    if (threadIdx.x == 0) {
      // the block's first particle sets b0 (a particle outside the grid
      // there leaves the whole block on the direct path)
      int ok = 0;
      int64_t bb = 0;
      const int64_t qf = (int64_t)blockIdx.x * kT * kDepChunk;
      if (qf < n && S <= kDepLdsSpecies) {
        const int i = p.c[0][qf], j = p.c[1][qf], k = p.c[2][qf];
        if (!(i < 0 || j < 0 || k < 0 || i > g.nx || j > g.ny || k > g.nz)) {
          ok = 1;
          bb = (int64_t)i + X * j + XY * k;
        }
      }
      seg_on = ok;
      seg_b0 = bb;
    }
    for (int e = threadIdx.x; e < 4 * kDepSeg * kDepLdsSpecies; e += kT) seg[e] = 0;
    __syncthreads();
    on = seg_on != 0;
    b0 = seg_b0;
  }
  int64_t cur_base = -1;
  int cur_s = -1;
  unsigned long long W[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long Wf[FLUX ? 24 : 1] = {};
  auto flush = [&]() {
    if (cur_base < 0) return;
    if (!FLUX && on && cur_base >= b0 && cur_base - b0 + 1 < kDepSeg) {
      const int off = (int)(cur_base - b0);
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        atomicAdd(&seg[((o >> 1) * kDepSeg + off + (o & 1)) * kDepLdsSpecies + cur_s], W[o]);
        W[o] = 0;
      }
      return;
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int64_t node = cur_base + (o & 1) + X * ((o >> 1) & 1) + XY * (o >> 2);
      atomicAdd(&Nfix[node * S + cur_s], W[o]);
      W[o] = 0;
      if (FLUX) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          atomicAdd(&Ffix[(node * 3 + d) * S + cur_s], Wf[3 * o + d]);
          Wf[3 * o + d] = 0;
        }
      }
    }
  };
  int nerr = 0;
  for (int64_t q = q0; q < q1; ++q) {
    const int i = p.c[0][q], j = p.c[1][q], k = p.c[2][q];
    if (i < 0 || j < 0 || k < 0 || i > g.nx || j > g.ny || k > g.nz) {
      ++nerr;
      continue;
    }
    const int s = p.sp[q] - 1;
    const int64_t base = (int64_t)i + X * j + XY * k;
    if (base != cur_base || s != cur_s) {
      flush();
      cur_base = base;
      cur_s = s;
    }
    const float fx = p.f[0][q], fy = p.f[1][q], fz = p.f[2][q];
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (FLUX) {
      v[0] = p.v[0][q];
      v[1] = p.v[1][q];
      v[2] = p.v[2][q];
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const float w = wx[o & 1] * wy[(o >> 1) & 1] * wz[o >> 2];
      W[o] += (unsigned long long)(double(w) * kFix);
      if (FLUX) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          Wf[3 * o + d] += (unsigned long long)llrint((double)w * (double)v[d] * kFluxFix);
      }
    }
  }
  flush();
  if (nerr) atomicAdd(errors, nerr);
  if constexpr (!FLUX) {
    __syncthreads();
    if (on) {
      for (int e = threadIdx.x; e < 4 * kDepSeg * kDepLdsSpecies; e += kT) {
        const unsigned long long v = seg[e];
        if (!v) continue;
        const int sidx = e % kDepLdsSpecies, r = e / kDepLdsSpecies;
        const int sg_ = r / kDepSeg, off = r % kDepSeg;
        const int64_t node = b0 + off + X * (sg_ & 1) + XY * (sg_ >> 1);
        atomicAdd(&Nfix[node * S + sidx], v);
      }
    }
  }
}

// Integer guard folds (ParticlesGridify.f90:129-285, z, y, x) and
// N = W * 2^-40 * n2p, C = sum_s N * charge.
__device__ __forceinline__ int sp_fold_src(int m, int n, int flow, int src[2], bool *zero) {
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

__global__ void __launch_bounds__(kT) k_sp_fold(GridDims g, const unsigned long long *__restrict__ Nfix,
                                                 double *__restrict__ N, double *__restrict__ C,
                                                 const double *consts) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int sx[2], sy[2], sz[2];
  bool zx, zy, zz;
  const int nxs = sp_fold_src(i, g.nx, g.flow[0], sx, &zx);
  const int nys = sp_fold_src(j, g.ny, g.flow[1], sy, &zy);
  const int nzs = sp_fold_src(k, g.nz, g.flow[2], sz, &zz);
  double cacc = 0.0;
  for (int s = 0; s < g.S; ++s) {
    unsigned long long w = 0;
    if (!zx && !zy && !zz)
      for (int a = 0; a < nxs; ++a)
        for (int b = 0; b < nys; ++b)
          for (int c = 0; c < nzs; ++c) w += Nfix[node_index(g, sx[a], sy[b], sz[c]) * g.S + s];
    const double v = (double)w * kFixInv * consts[s];
    N[node * g.S + s] = v;
    cacc = cacc + v * consts[8 + s];
  }
  C[node] = cacc;
}

// Integer guard folds of the fluxes, F = Wf * 2^-24 * n2p (ParticlesGridifyAll.f90:238-420).
__global__ void __launch_bounds__(kT) k_sp_fold_flux(GridDims g,
                                                      const unsigned long long *__restrict__ Ffix,
                                                      double *__restrict__ F, const double *consts) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int sx[2], sy[2], sz[2];
  bool zx, zy, zz;
  const int nxs = sp_fold_src(i, g.nx, g.flow[0], sx, &zx);
  const int nys = sp_fold_src(j, g.ny, g.flow[1], sy, &zy);
  const int nzs = sp_fold_src(k, g.nz, g.flow[2], sz, &zz);
  const int nc = 3 * g.S;
  for (int q = 0; q < nc; ++q) {
    unsigned long long w = 0;
    if (!zx && !zy && !zz)
      for (int a = 0; a < nxs; ++a)
        for (int b = 0; b < nys; ++b)
          for (int c = 0; c < nzs; ++c) w += Ffix[node_index(g, sx[a], sy[b], sz[c]) * nc + q];
    F[node * nc + q] = (double)(long long)w * kFluxFixInv * consts[q % g.S];
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision Jacobi

struct MixCoef {
  double ax, ay, az, af, dx2, dy2, dz2;
  float axf, ayf, azf, aff, idx2f, idy2f, idz2f;
};



// neighbours of interior cell c with periodic wrap (flow guards read as is)
struct Nb {
  int64_t e, w, n, s, t, b;
};

__device__ __forceinline__ Nb neighbours(const GridDims &g, int i, int j, int k, int64_t c) {
  const int64_t X = g.X, XY = (int64_t)g.X * g.Y;
  Nb r;
  r.e = (!g.flow[0] && i == g.nx + 1) ? c - (int64_t)(g.nx - 1) : c + 1;
  r.w = (!g.flow[0] && i == 2) ? c + (int64_t)(g.nx - 1) : c - 1;
  r.n = (!g.flow[1] && j == g.ny + 1) ? c - (int64_t)(g.ny - 1) * X : c + X;
  r.s = (!g.flow[1] && j == 2) ? c + (int64_t)(g.ny - 1) * X : c - X;
  r.t = (!g.flow[2] && k == g.nz + 1) ? c - (int64_t)(g.nz - 1) * XY : c + XY;
  r.b = (!g.flow[2] && k == 2) ? c + (int64_t)(g.nz - 1) * XY : c - XY;
  return r;
}

__device__ __forceinline__ double sum_partials(const double *__restrict__ P, int n, double *sh) {
  double v = 0.0;
  const int nt = (int)(blockDim.x * blockDim.y * blockDim.z);
  for (int q = lin_tid(); q < n; q += nt) v = v + P[q];
  return block_sum(v, sh);
}

// One launch of the mixed-precision solve.  Launch number `seq` reads the
// stop state of launch seq-1 and writes its own (block 0).  If `check` >= 1,
// the `nin` partials `Pin` of the previous launch hold the residual of
// iterate `check`, and the stop test of loop iteration check+1 is applied to
// it (CalcEpotential.f90:60-62); every block evaluates the same fixed-order
// sum, so all reach the same decision without atomics.
struct MxLaunch {
  int32_t seq;
  int32_t base;          // iterate held by u
  int32_t j;             // inner sweep: e_j -> e_{j+1}
  int32_t check;         // iterate whose residual Pin holds, or -1
  int32_t m;             // inner sweeps of this outer step (update launch)
  int32_t nin;           // partials in Pin
  int32_t pairs;         // the solve runs launches of more than one sweep
  int32_t test;          // two-sweep launch: evaluate the stop-test residual of e_j
  int32_t prevtest;      // the previous launch wrote partials whatever its near flag
  int32_t ibuf;          // correction buffer holding e_j (update launch: e_m)
  int32_t cbuf;          // correction buffer holding iterate `check`
};

// Compact FP32 layout of the correction, its right-hand side and the
// boundary-plane copies: interior cells only, x fastest, rows padded with
// zeros to a multiple of 4 cells so that a lane moves 4 cells with one
// 16-byte access.  Cell (x, y, z), 0-based, sits at x + NX y + P z.
struct CompactDims {
  uint32_t NX, P;        // row stride (nx rounded up to 4), plane stride NX * ny
};

// compact index of interior node (i, j, k), 1-based (2..n+1)
__device__ __forceinline__ uint32_t compact_index(const CompactDims &cd, int i, int j, int k) {
  return (uint32_t)(i - 2) + cd.NX * (uint32_t)(j - 2) + cd.P * (uint32_t)(k - 2);
}

constexpr int kSwThreads = 256;           // threads per block of the sweep kernel
constexpr int kSwRows = kSwThreads / 64;  // rows of a sweep tile, one wave each
constexpr int kOutRows = kStencilT / 64;  // rows of an outer-step tile, one wave each
constexpr int kSwSeg = 256;               // cells of a row segment: 64 lanes x 4

// Uniform geometry of the sweep kernel, kept small (32-bit, no FP64 box
// data) so it stays in scalar registers.
struct SweepGeo {
  int32_t nx, ny, nz;
  uint32_t NX, P;        // compact strides
  int32_t px, py, pz;    // periodic flags
  int32_t ntx, nty, nzc; // tiles: row segments, row groups, z chunks of an XCD slab
  int32_t TP, W;         // tiles per XCD slab, blocks per XCD
};

struct SweepCoef {
  float wx, wy, wz, kappa, af, idx2, idy2, idz2;
};

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }

// One cell of the FP32 inner sweep, CalcEpotential.f90:74-80 on the
// correction: x-neighbours ee/ew, y n/s, z t/b, centre c, right-hand side r.
// The weights sum to 1/2 (kappa carries what rounding leaves, zero on
// uniform grids): both ends of the Jacobi spectrum, whose modes decay by
// ~1e-5 per sweep at 256^3, stay where the FP64 iteration has them.  The
// terms are summed left to right with fused multiply-adds (8 operations
// against 12 with every product rounded), in one fixed order so that every
// kernel that sweeps gives the same bits.  At 256^3 the fused form makes the
// two-sweep kernel 2-3% faster; the solve of a step then stops 30-40 sweeps
// ahead of the FP64 solve instead of 8-13, and phi is within 4.8e-6 of the
// FP64 form's instead of 2.7e-6.
__device__ __forceinline__ float jac_cell(const SweepCoef &cf, float ee, float ew, float n, float s,
                                          float t, float b, float c, float r) {
  float v = cf.wx * (ee + ew);
  v = __builtin_fmaf(cf.wy, n + s, v);
  v = __builtin_fmaf(cf.wz, t + b, v);
  v = __builtin_fmaf(cf.kappa, c, v);
  return __builtin_fmaf(-cf.af, r, v);
}

// The stop test of iterate L.check (CalcEpotential.f90:60-61) on the
// previous launch's partials, when that launch tested: the host marked it
// (prevtest) or its near flag was set (it tested every launch then).  Each
// test also sets `near`:
// whether the residual, extrapolated at the decay per sweep seen since the
// previous test to the next sparse test (kNearSpan sweeps on), is within
// kNear of the threshold.  While near, the two-sweep kernel tests every
// launch, so the solve stops at most a sweep after the first iterate that
// passes; the flag is taken again at every test, so a fast early decay that
// overstates the next one is corrected by the tests that follow.
constexpr int kPairTestEvery = 16;  // two-sweep launches per stop test, far from the stop
constexpr double kNear = 1.25;
constexpr double kNearSpan = 2.0 * kPairTestEvery;   // sweeps between sparse tests

__device__ __forceinline__ StopSlot mx_decide(JacobiState *st, const MxLaunch &L,
                                              const double *__restrict__ Pin, double *sh,
                                              int *decide) {
  const StopSlot prev = st->slot[(L.seq - 1) & 1];
  StopSlot cur = prev;
  if (!prev.done && L.check >= 1 && (L.prevtest || prev.near)) {
    const double S = sum_partials(Pin, L.nin, sh);
    const bool tested = S >= 0.0;
    const double rn = sqrt(tested ? S : 0.0);
    if (lin_tid() == 0) {
      double next = rn;
      const int64_t gap = (int64_t)L.check - prev.itlast;
      if (prev.rlast > 0.0 && rn > 0.0 && gap > 0)
        next = rn * pow(rn / prev.rlast, kNearSpan / (double)gap);
      *decide = ((tested && L.check + 1 <= st->maxit && rn <= st->thresh) ? 1 : 0) |
                ((tested && fmin(rn, next) <= kNear * st->thresh) ? 2 : 0);
    }
    __syncthreads();
    if (*decide & 1) {
      cur.done = 1;
      cur.result = L.check;
      cur.buf = L.cbuf;
    }
    if (tested) {
      cur.near = (*decide & 2) ? 1 : 0;
      cur.rlast = rn;
      cur.itlast = L.check;
    }
  }
  return cur;
}

// Geometry of the outer-step kernel: guard-layout FP64 strides, compact FP32
// strides, and a schedule of 64-cell row segments (one cell per lane) x
// 4-row groups x z chunks of an XCD slab.
struct OuterGeo {
  int32_t nx, ny, nz;
  uint32_t X, XY;        // guard-layout strides
  uint32_t NX, P;        // compact strides
  int32_t px, py, pz;    // periodic flags
  int32_t ntx, nty, nzc, TP, W;
};

// Outer step of the mixed solve.  u' = u + e_m, the correction the inner
// sweeps of the previous outer step left in `em` (UPD; the first outer step
// has none), goes to the other FP64 buffer `ud` (interior nodes; the guards
// of both buffers hold the same boundary values); r = b - A u' in FP64 with
// fresh periodic images, r32 = r (the right-hand side of the next inner
// sweeps); partials of |r_stale|^2 for the stop test of iterate `base`.  The
// reference evaluates that residual with the periodic halo of the previous
// iterate (CalcEpotential.f90:83-92), so a neighbour across a periodic
// boundary takes U_{base-1} = u' - (e_m - e_{m-1}); eo holds e_{m-1}.
template <bool UPD>
__global__ void __launch_bounds__(kStencilT, 4) k_mx_outer(OuterGeo og, double dx2, double dy2,
                                                   double dz2, const double *__restrict__ us,
                                                   double *__restrict__ ud,
                                                   const double *__restrict__ b,
                                                   float *__restrict__ r32,
                                                   const float *__restrict__ em,
                                                   const float *__restrict__ eo,
                                                   const double *__restrict__ Pin,
                                                   double *__restrict__ Pout, JacobiState *st,
                                                   MxLaunch L) {
  __shared__ double sh[kStencilT / 64];
  __shared__ int decide;
  const StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  if (cur.done) return;
  const int nx = og.nx, ny = og.ny, nz = og.nz;
  const uint32_t X = og.X, XY = og.XY, NX = og.NX, P = og.P;
  const int lane = (int)threadIdx.x;
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  for (int t = w; t < og.TP; t += og.W) {
    const int zi = t % og.nzc, tt = t / og.nzc;
    const int tx = tt % og.ntx, ty = tt / og.ntx;
    const int ka = z0 + ((z1 - z0) * zi) / og.nzc, kb = z0 + ((z1 - z0) * (zi + 1)) / og.nzc;
    const int j = ty * kOutRows + (int)threadIdx.y;
    if (j >= ny || kb <= ka) continue;                      // the whole wave
    const int i0 = tx * 64 + lane;
    const bool live = i0 < nx;                              // lanes past the row end store nothing
    const int i = live ? i0 : nx - 1;
    // node of (i, j, plane 0 of the guard layout) and compact cell (i, j, 0)
    const uint32_t g0 = (uint32_t)(i + 1) + X * (uint32_t)(j + 1);
    const uint32_t c0 = (uint32_t)i + NX * (uint32_t)j;
    // neighbour offsets: node, compact cell, and whether the neighbour is an
    // interior cell (periodic image) or a flow guard (boundary value, no e)
    const bool wE = og.px && i == nx - 1, wW = og.px && i == 0;     // wrapped
    const bool gE = !og.px && i == nx - 1, gW = !og.px && i == 0;   // flow guard
    const uint32_t dgE = wE ? 0u - (uint32_t)(nx - 1) : 1u, dcE = wE ? 0u - (uint32_t)(nx - 1) : 1u;
    const uint32_t dgW = wW ? (uint32_t)(nx - 1) : 0u - 1u, dcW = wW ? (uint32_t)(nx - 1) : 0u - 1u;
    const bool wN = og.py && j == ny - 1, wS = og.py && j == 0;
    const bool gN = !og.py && j == ny - 1, gS = !og.py && j == 0;
    const uint32_t dgN = wN ? 0u - (uint32_t)(ny - 1) * X : X;
    const uint32_t dcN = wN ? 0u - (uint32_t)(ny - 1) * NX : NX;
    const uint32_t dgS = wS ? (uint32_t)(ny - 1) * X : 0u - X;
    const uint32_t dcS = wS ? (uint32_t)(ny - 1) * NX : 0u - NX;
    // x-neighbours come from the adjacent lanes, except at the segment ends
    // and the domain boundary, where they are read
    const bool xe_mem = lane == 63 || i == nx - 1, xw_mem = lane == 0 || i == 0;
    // u and e on plane kk of the z-window, -1 <= kk <= nz (past the ends: the
    // periodic image, or the guard's boundary value with e = 0); u' = u + e
    struct Zv {
      double u;
      float e;
    };
    auto zval = [&](int kk) -> Zv {
      Zv z;
      z.e = 0.0f;
      const bool guard = !og.pz && (kk < 0 || kk >= nz);
      const int kw = guard ? kk : (kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk));
      if (UPD && !guard) z.e = em[c0 + P * (uint32_t)kw];
      z.u = us[g0 + XY * (uint32_t)(kw + 1)];
      return z;
    };
    // the rest of plane kk's inputs; loads that may be skipped go first (see
    // the sweep kernel); e = 0 where a flow guard is read
    struct OIn {
      double un, us, b, uxe, uxw;
      float en, es, exe, exw;
      float qxe, qxw, qn, qs, qz;    // e_{m-1} at the neighbours across periodic boundaries
    };
    auto plane_in = [&](int kk) -> OIn {
      const uint32_t gk = g0 + XY * (uint32_t)(kk + 1), ck = c0 + P * (uint32_t)kk;
      OIn a;
      a.uxe = 0.0;
      a.uxw = 0.0;
      a.en = 0.0f;
      a.es = 0.0f;
      a.exe = 0.0f;
      a.exw = 0.0f;
      a.qxe = 0.0f;
      a.qxw = 0.0f;
      a.qn = 0.0f;
      a.qs = 0.0f;
      a.qz = 0.0f;
      if (UPD) {
        if (wE) a.qxe = eo[ck + dcE];
        if (wW) a.qxw = eo[ck + dcW];
        if (wN) a.qn = eo[ck + dcN];
        if (wS) a.qs = eo[ck + dcS];
        if (og.pz && kk == nz - 1) a.qz = eo[ck - P * (uint32_t)(nz - 1)];
        else if (og.pz && kk == 0) a.qz = eo[ck + P * (uint32_t)(nz - 1)];
        if (xe_mem && !gE) a.exe = em[ck + dcE];
        if (xw_mem && !gW) a.exw = em[ck + dcW];
        if (!gN) a.en = em[ck + dcN];
        if (!gS) a.es = em[ck + dcS];
      }
      if (xe_mem) a.uxe = us[gk + dgE];
      if (xw_mem) a.uxw = us[gk + dgW];
      a.un = us[gk + dgN];
      a.us = us[gk + dgS];
      a.b = b[gk];
      return a;
    };
    // plane k from its inputs and the z-window (zb, zc, zt) = planes k-1, k, k+1
    auto do_plane = [&](int k, const OIn &a, const Zv &zb, const Zv &zc, const Zv &zt) {
      const uint32_t gk = g0 + XY * (uint32_t)(k + 1), ck = c0 + P * (uint32_t)k;
      const double uc = UPD ? zc.u + (double)zc.e : zc.u;
      const double ut = UPD ? zt.u + (double)zt.e : zt.u;
      const double ub = UPD ? zb.u + (double)zb.e : zb.u;
      double ue = __shfl(uc, (lane + 1) & 63, 64), uw = __shfl(uc, (lane + 63) & 63, 64);
      if (xe_mem) ue = UPD ? a.uxe + (double)a.exe : a.uxe;
      if (xw_mem) uw = UPD ? a.uxw + (double)a.exw : a.uxw;
      const double un = UPD ? a.un + (double)a.en : a.un;
      const double us_ = UPD ? a.us + (double)a.es : a.us;
      const double u2 = 2.0 * uc;
      const double r = a.b - ((ue + uw - u2) / dx2 + (un + us_ - u2) / dy2 + (ut + ub - u2) / dz2);
      double rs = r;
      const bool wT = og.pz && k == nz - 1, wB = og.pz && k == 0;
      if (UPD && (wE || wW || wN || wS || wT || wB)) {
        const double se = wE ? ue - ((double)a.exe - (double)a.qxe) : ue;
        const double sw = wW ? uw - ((double)a.exw - (double)a.qxw) : uw;
        const double sn = wN ? un - ((double)a.en - (double)a.qn) : un;
        const double ss = wS ? us_ - ((double)a.es - (double)a.qs) : us_;
        const double st_ = wT ? ut - ((double)zt.e - (double)a.qz) : ut;
        const double sb = wB ? ub - ((double)zb.e - (double)a.qz) : ub;
        rs = a.b - ((se + sw - u2) / dx2 + (sn + ss - u2) / dy2 + (st_ + sb - u2) / dz2);
      }
      if (live) {
        ud[gk] = uc;
        r32[ck] = (float)r;
        r2 = r2 + rs * rs;
      }
    };
    // four planes per pass, as in the sweep kernel
    auto zl = [&](int kk) { return zval(kk < kb ? kk : kb); };
    auto pl = [&](int kk) { return plane_in(kk < kb ? kk : kb - 1); };
    Zv z0 = zl(ka - 1), z1 = zl(ka), z2 = zl(ka + 1), z3 = zl(ka + 2);
    OIn A = pl(ka), B = pl(ka + 1);
    for (int k = ka; k < kb; k += 4) {
      do_plane(k, A, z0, z1, z2);
      if (k + 1 >= kb) break;
      z0 = zl(k + 3);
      A = pl(k + 2);
      do_plane(k + 1, B, z1, z2, z3);
      if (k + 2 >= kb) break;
      z1 = zl(k + 4);
      B = pl(k + 3);
      do_plane(k + 2, A, z2, z3, z0);
      if (k + 3 >= kb) break;
      z2 = zl(k + 5);
      A = pl(k + 4);
      do_plane(k + 3, B, z3, z0, z1);
      z3 = zl(k + 6);
      B = pl(k + 5);
    }
  }
  const double sblk = block_sum(r2, sh);
  if (lin_tid() == 0) Pout[blockIdx.x] = sblk;
}

// Inner FP32 sweep e_j -> e_{j+1} (e_0 = 0; e_k lives in E[k % 2]) and, for
// j >= 1, partials of the residual of iterate base+j, r - A e_j, accumulated
// in FP64.  Flow guards of e are zero: the correction keeps the boundary.
// Two buffers suffice: the stop test of iterate base+j runs at the start of
// the launch that would overwrite e_j, before any block writes.
//
// A wave owns one row segment of 256 cells, 4 per lane, and marches up a z
// chunk of its XCD's slab; x-neighbours inside the segment come from the
// adjacent lanes.  A flow guard, and the z-window of lanes past the row end,
// read the zero row stored after the last plane of E, so that no loaded value
// needs a select; padding cells past x = nx - 1 are kept at zero, which makes
// them the flow guard of x = nx - 1.  PX: x periodic; PAD: nx is not a
// multiple of 4.
//
// The residual is that of the stop test, which the reference evaluates with
// the periodic halo of the previous iterate (CalcEpotential.f90:83-92): a
// neighbour across a periodic boundary takes e_{j-1}.  e_{j-1} is kept, for
// the cells of the periodic boundary planes, in slot (j-1) % 3 of the
// boundary copies (`eo_prev`, null for j < 2); this launch leaves e_{j+1}
// there in slot (j+1) % 3 (`eo_out`).  e is e_j (null for j = 0) and out
// receives e_{j+1}.
template <bool PX, bool PAD>
__global__ void __launch_bounds__(kSwThreads, 1024 / kSwThreads) k_mx_sweep(SweepGeo sg, SweepCoef cf,
                                                   const float *__restrict__ r32,
                                                   const float *__restrict__ e,
                                                   float *__restrict__ out,
                                                   float *__restrict__ eo_out,
                                                   const float *__restrict__ eo_prev,
                                                   const double *__restrict__ Pin,
                                                   double *__restrict__ Pout, JacobiState *st,
                                                   MxLaunch L,
    unsigned long long *ks) {
  __shared__ double sh[kSwThreads / 64];
  __shared__ int decide;
  kstamp_begin(ks);
  StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (!cur.done && (int64_t)L.base + L.j >= st->last) {   // iterate 2*maxit+1 reached (:58)
    cur.done = 1;
    cur.result = st->last;
    cur.buf = L.ibuf;
  }
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }
  const bool res = L.j >= 1;
  const uint32_t NX = sg.NX, P = sg.P;
  const int nx = sg.nx, ny = sg.ny, nz = sg.nz;
  const int lane = (int)threadIdx.x;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t zrow = P * (uint32_t)nz;                   // the zero row of E
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  for (int t = w; t < sg.TP; t += sg.W) {
    const int zi = t % sg.nzc, tt = t / sg.nzc;
    const int tx = tt % sg.ntx, ty = tt / sg.ntx;
    const int ka = z0 + ((z1 - z0) * zi) / sg.nzc, kb = z0 + ((z1 - z0) * (zi + 1)) / sg.nzc;
    const int j = ty * kSwRows + (int)threadIdx.y;
    if (j >= ny || kb <= ka) continue;                      // the whole wave
    const int xs = tx * kSwSeg + 4 * lane;                  // the lane's first cell
    const bool live = xs < nx;                              // lanes past the row end store nothing
    const uint32_t x0 = live ? (uint32_t)xs : 0u;
    const uint32_t rowj = NX * (uint32_t)j, row = x0 + rowj;
    const bool zn = !sg.py && j == ny - 1, zs = !sg.py && j == 0;   // flow guards
    const uint32_t rown = x0 + NX * (uint32_t)(j == ny - 1 ? 0 : j + 1);
    const uint32_t rows = x0 + NX * (uint32_t)(j == 0 ? ny - 1 : j - 1);
    const int ql = nx - 1 - xs;                             // x = nx - 1 at cell ql
    const bool hasf = PX && xs == 0;                        // the lane holds x = 0 (periodic x)
    const bool hasl = PX && live && ql < 4;                 // ... x = nx - 1
    const bool yN = sg.py && j == ny - 1, yS = sg.py && j == 0;
    const bool edge_row = hasf || hasl || yN || yS;
    auto inrow = [&](int q) { return live && (!PAD || q <= ql); };
    if (!e) {
      for (int k = ka; k < kb; ++k) {
        const uint32_t c = row + P * (uint32_t)k;
        const float4 rc = ld4(r32 + c);
        float4 v;
        v.x = inrow(0) ? -cf.af * rc.x : 0.0f;
        v.y = inrow(1) ? -cf.af * rc.y : 0.0f;
        v.z = inrow(2) ? -cf.af * rc.z : 0.0f;
        v.w = inrow(3) ? -cf.af * rc.w : 0.0f;
        if (live) {
          st4(out + c, v);
          if (edge_row || (sg.pz && (k == 0 || k == nz - 1))) st4(eo_out + c, v);
        }
      }
      continue;
    }
    // e_j on plane kk, -1 <= kk <= nz (past the ends: the periodic image or
    // the zero row)
    auto zval = [&](int kk) -> float4 {
      const int kw = kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk);
      const bool zero = !live || (!sg.pz && (kk < 0 || kk >= nz));
      return ld4(e + (zero ? zrow + x0 : row + P * (uint32_t)kw));
    };
    // what plane kk needs besides the z-window: y-neighbours, right-hand
    // side, the x-neighbours no lane of the wave holds (lane 0: the previous
    // segment's last cell or the image of x = nx - 1; lane 63: the next
    // segment's first cell; the lane with x = nx - 1: the image of x = 0), and
    // e_{j-1} (0 for j = 1) at the neighbours across periodic boundaries
    struct PlaneIn {
      float4 n, s, r, py, pz;
      float w, eseg, ewrap, pw, pe;
    };
    // Loads that may be skipped go first: a wait for this group's last load
    // then also covers them, and the waits of later groups can count the
    // unconditional loads exactly.
    auto plane_in = [&](int kk) -> PlaneIn {
      const uint32_t pk = P * (uint32_t)kk;
      PlaneIn a;
      a.w = 0.0f;
      a.eseg = 0.0f;
      a.ewrap = 0.0f;
      a.py = zero4;
      a.pz = zero4;
      a.pw = 0.0f;
      a.pe = 0.0f;
      if (eo_prev) {
        if (yN) a.py = ld4(eo_prev + x0 + pk);                                   // row 0
        else if (yS) a.py = ld4(eo_prev + x0 + NX * (uint32_t)(ny - 1) + pk);    // row ny - 1
        if (sg.pz && kk == nz - 1) a.pz = ld4(eo_prev + row);                    // plane 0
        else if (sg.pz && kk == 0) a.pz = ld4(eo_prev + row + P * (uint32_t)(nz - 1));
        if (hasf) a.pw = eo_prev[rowj + pk + (uint32_t)(nx - 1)];
        if (hasl) a.pe = eo_prev[rowj + pk];
      }
      if (sg.ntx > 1) {
        if (lane == 0 && xs > 0) a.w = e[row + pk - 1u];
        if (lane == 63 && xs + 4 < nx) a.eseg = e[row + pk + 4u];
      }
      if (hasf) a.w = e[rowj + pk + (uint32_t)(nx - 1)];
      if (hasl) a.ewrap = e[rowj + pk];
      a.n = ld4(e + (zn ? zrow + x0 : rown + pk));
      a.s = ld4(e + (zs ? zrow + x0 : rows + pk));
      a.r = ld4(r32 + row + pk);
      return a;
    };
    // plane k from its inputs and the z-window (zb, zc, zt) = planes k-1, k, k+1
    auto do_plane = [&](int k, const PlaneIn &a, const float4 &zb, const float4 &zc,
                        const float4 &zt) {
      const uint32_t c = row + P * (uint32_t)k;
      float wv = __shfl(zc.w, (lane + 63) & 63, 64);
      float ev = __shfl(zc.x, (lane + 1) & 63, 64);
      if (lane == 0) wv = a.w;
      if (lane == 63) ev = a.eseg;
      const bool zT = sg.pz && k == nz - 1, zB = sg.pz && k == 0;
      const float ce[4] = {zc.x, zc.y, zc.z, zc.w};
      const float cn[4] = {a.n.x, a.n.y, a.n.z, a.n.w}, cs[4] = {a.s.x, a.s.y, a.s.z, a.s.w};
      const float ct[4] = {zt.x, zt.y, zt.z, zt.w}, cb[4] = {zb.x, zb.y, zb.z, zb.w};
      const float cr[4] = {a.r.x, a.r.y, a.r.z, a.r.w};
      const float qy[4] = {a.py.x, a.py.y, a.py.z, a.py.w};
      const float qz[4] = {a.pz.x, a.pz.y, a.pz.z, a.pz.w};
      float v[4], sq = 0.0f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float ew = q == 0 ? wv : ce[q - 1];
        float ee = q == 3 ? ev : ce[q + 1];
        if (hasl && q == ql) ee = a.ewrap;
        v[q] = jac_cell(cf, ee, ew, cn[q], cs[q], ct[q], cb[q], ce[q], cr[q]);
        if (PAD && q > ql) v[q] = 0.0f;
        // stop-test residual: e_{j-1} across the periodic boundaries
        const float se = (hasl && q == ql) ? a.pe : ee;
        const float sw = (hasf && q == 0) ? a.pw : ew;
        const float sn = yN ? qy[q] : cn[q];
        const float ss = yS ? qy[q] : cs[q];
        const float st_ = zT ? qz[q] : ct[q];
        const float sb = zB ? qz[q] : cb[q];
        const float e2 = 2.0f * ce[q];
        const float rr = cr[q] - ((se + sw - e2) * cf.idx2 + (sn + ss - e2) * cf.idy2 +
                                  (st_ + sb - e2) * cf.idz2);
        if (inrow(q)) sq = sq + rr * rr;       // the lane's 4 squares in FP32
      }
      r2 = r2 + (double)sq;
      const float4 v4 = make_float4(v[0], v[1], v[2], v[3]);
      if (live) {
        st4(out + c, v4);
        if (edge_row || zT || zB) st4(eo_out + c, v4);
      }
    };
    // Four planes per pass: the window registers z0..z3 rotate by name, and
    // each load goes out two planes before its use, right after the plane
    // that last read its register.
    auto zl = [&](int kk) { return zval(kk < kb ? kk : kb); };
    auto pl = [&](int kk) { return plane_in(kk < kb ? kk : kb - 1); };
    float4 z0 = zl(ka - 1), z1 = zl(ka), z2 = zl(ka + 1), z3 = zl(ka + 2);
    PlaneIn A = pl(ka), B = pl(ka + 1);
    for (int k = ka; k < kb; k += 4) {
      do_plane(k, A, z0, z1, z2);
      if (k + 1 >= kb) break;
      z0 = zl(k + 3);
      A = pl(k + 2);
      do_plane(k + 1, B, z1, z2, z3);
      if (k + 2 >= kb) break;
      z1 = zl(k + 4);
      B = pl(k + 3);
      do_plane(k + 2, A, z2, z3, z0);
      if (k + 3 >= kb) break;
      z2 = zl(k + 5);
      A = pl(k + 4);
      do_plane(k + 3, B, z3, z0, z1);
      z3 = zl(k + 6);
      B = pl(k + 5);
    }
  }
  const double sblk = block_sum(r2, sh);
  if (lin_tid() == 0 && res) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// Two inner sweeps per launch, e_j -> e_{j+2}, for rows of one segment
// (nx <= 256).  Each wave owns R consecutive rows y0..y0+R-1 and marches up a
// z chunk as in k_mx_sweep; at every plane step p it computes the first
// sweep, e_{j+1}, on rows y0-1..y0+R at plane p (its rows and one more on
// each side, which the neighbouring waves compute as well, rather than
// sharing them through LDS with a barrier per plane), and the second sweep,
// e_{j+2}, on its rows at plane p-1 from the e_{j+1} planes p-2..p it holds.
// The unique traffic per launch is that of one sweep (e_j and r32 read,
// e_{j+2} written); rows of e_j shared by neighbouring waves come mostly from
// cache.  Every cell's arithmetic is k_mx_sweep's, so e_{j+2} is bit for bit
// what two launches of it give.  The stop-test residual is that of e_j (taken
// from the first sweep of the wave's rows, with the stale halo of e_{j-1}),
// evaluated in the launches the host marks and in every launch while the last
// test found the stop near (mx_decide); the boundary copies of e_{j+1} go to
// `eo_out`, for the residual of e_{j+2} in the next launch.  x-neighbours
// come from DPP wave shifts and the periodic image of x from readlane; no
// load depends on the lane beyond its address, so the wave stays convergent
// where it exchanges values.  FIRST: j = 0 (e_0 = 0), the first sweep is
// -af r and nothing of e is read.
//
// MULTI: rows wider than one segment (x not periodic).  Segment k keeps the
// cells [kPairKeep*k, kPairKeep*k + kPairKeep) and its wave holds the 256
// cells from four before them: the first sweep of the cells on either side
// of the kept ones is then the wave's own (lanes 0 and 63, whose outer cells
// lack a neighbour, feed only cells that are not kept), at the price of
// sweeping 256 cells for 248.
constexpr int kPairRows = 2;     // rows per wave of the two-sweep kernel
constexpr int kPairKeep = 248;   // cells a wave keeps of a segment (MULTI)

template <bool PX, bool PAD, int R, bool FIRST, bool MULTI>
__global__ void __launch_bounds__(kSwThreads, 2) k_mx_pair(SweepGeo sg, SweepCoef cf,
                                                  const float *__restrict__ r32,
                                                  const float *__restrict__ e,
                                                  float *__restrict__ out,
                                                  float *__restrict__ eo_out,
                                                  const float *__restrict__ eo_prev,
                                                  const double *__restrict__ Pin,
                                                  double *__restrict__ Pout, JacobiState *st,
                                                  MxLaunch L,
    unsigned long long *ks) {
  static_assert(!(MULTI && PX), "overlapping segments need x not periodic");
  __shared__ double sh[kSwThreads / 64];
  __shared__ int decide;
  kstamp_begin(ks);
  StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (!cur.done && (int64_t)L.base + L.j >= st->last) {   // iterate 2*maxit+1 reached (:58)
    cur.done = 1;
    cur.result = st->last;
    cur.buf = L.ibuf;
  }
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }
  // the stop-test residual of e_j: in the launches the host marks, and in
  // every launch while the last test found the stop near
  const bool res = !FIRST && (L.test || cur.near);
  constexpr bool first = FIRST;
  const uint32_t NX = sg.NX, P = sg.P;
  const int nx = sg.nx, ny = sg.ny, nz = sg.nz;
  const int lane = (int)threadIdx.x;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t zrow = P * (uint32_t)nz;                   // the zero row of E
  const int Lw = (nx - 1) >> 2, qw = (nx - 1) & 3;          // lane and cell of x = nx - 1
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  auto shr = [](float v) {      // from the lane below (0 into lane 0)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
  };
  auto shl = [](float v) {      // from the lane above (0 into lane 63)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
  };
  auto bcast = [](float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  };
  // periodic image of plane pp, -2 <= pp <= nz + 1 (also for nz = 1)
  auto wrapz = [&](int pp) {
    int kk = pp < 0 ? pp + nz : pp;
    kk = kk < 0 ? kk + nz : kk;
    kk = kk >= nz ? kk - nz : kk;
    return kk >= nz ? kk - nz : kk;
  };
  for (int t = w; t < sg.TP; t += sg.W) {
    const int zi = t % sg.nzc, tt = t / sg.nzc;
    const int tx = MULTI ? tt % sg.ntx : 0, ty = MULTI ? tt / sg.ntx : tt;
    const int ka = z0 + ((z1 - z0) * zi) / sg.nzc, kb = z0 + ((z1 - z0) * (zi + 1)) / sg.nzc;
    // the wave's first row (the same in every lane: scalar flags below)
    const int y0 = __builtin_amdgcn_readfirstlane((ty * kSwRows + (int)threadIdx.y) * R);
    if (y0 >= ny || kb <= ka) continue;                     // the whole wave
    const int xs = MULTI ? kPairKeep * tx - 4 + 4 * lane : 4 * lane;   // the lane's first cell
    const bool live = xs >= 0 && xs < nx;                   // lanes outside the row store nothing
    // the lane's cells are the wave's to store (MULTI: lanes 1..62)
    const bool keep = live && (!MULTI || (lane >= 1 && lane <= 62));
    const bool has0 = MULTI && xs == 0;                     // the lane holds x = 0 (flow guard west)
    const uint32_t x0 = live ? (uint32_t)xs : 0u;
    const int ql = nx - 1 - xs;                             // x = nx - 1 at cell ql
    const bool hasf = PX && xs == 0;                        // the lane holds x = 0 (periodic x)
    const bool hasl = PX && live && ql < 4;                 // ... x = nx - 1
    const bool haslast = live && ql < 4;                    // the lane holds x = nx - 1
    auto inrow = [&](int q) { return live && (!PAD || q <= ql); };
    // rows y0-2..y0+R+1 (index d: row y0+d-2): the lane's address of its
    // cells, whether the row is in the grid (or its periodic image), and
    // whether e_j is read there (else the zero row); flags are combined
    // without short-circuit so that every load is one select, not a branch
    uint32_t ra[R + 4];
    bool rv[R + 4], rok[R + 4];
#pragma unroll
    for (int d = 0; d < R + 4; ++d) {
      int yy = y0 + d - 2;
      bool ok = true;
      if (yy < 0 || yy >= ny) {
        if (sg.py) {
          yy = yy % ny;
          if (yy < 0) yy += ny;
        } else {
          ok = false;
          yy = y0;
        }
      }
      ra[d] = x0 + NX * (uint32_t)yy;
      rv[d] = ok;
      rok[d] = (!first) & live & ok;
    }
    // the wave's rows: in the grid (the last wave may hold fewer than R)
    bool own[R], yN[R], yS[R], edge[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      own[r] = y0 + r < ny;
      yN[r] = sg.py && y0 + r == ny - 1;
      yS[r] = sg.py && y0 + r == 0;
      edge[r] = hasf || hasl || yN[r] || yS[r];
    }
    const bool anyN = sg.py && y0 <= ny - 1 && ny - 1 < y0 + R;
    const bool anyS = sg.py && y0 == 0;
    const uint32_t zr = zrow + x0;
    auto pvalid = [&](int pp) { return sg.pz | ((pp >= 0) & (pp < nz)); };
    auto poff = [&](int pp) { return P * (uint32_t)wrapz(pp); };
    auto e0at = [&](int d, uint32_t o, bool v) -> float4 {
      if constexpr (FIRST) return zero4;       // e_0 = 0: nothing to read
      return ld4(e + ((v & rok[d]) ? ra[d] + o : zr));
    };
    // what plane step p reads: e_j on rows y0-1..y0+R at plane p+1 and on
    // rows y0-2, y0+R+1 at plane p, r32 on rows y0-1..y0+R at plane p, and
    // e_{j-1} at the neighbours of the wave's rows across periodic
    // boundaries (the stop test)
    struct StepIn {
      float4 f[2], r[R + 2], pyn, pys, pz[R];
      float pw[R], pe[R];
    };
    // loads into a (and the e_j rows at plane p+1 into cn) in place: the
    // stale values are read only at the steps and rows whose flags select
    // them, so they need no clearing per step
    auto step_in = [&](StepIn &a, float4 (&cn)[R + 2], int p) {
      if (res && p >= ka && p < kb) {
        const uint32_t pk = P * (uint32_t)p;
        if (anyN) a.pyn = ld4(eo_prev + x0 + pk);                               // row 0
        if (anyS) a.pys = ld4(eo_prev + x0 + NX * (uint32_t)(ny - 1) + pk);     // row ny - 1
        if (sg.pz && (p == nz - 1 || p == 0)) {
          const uint32_t pi = p == 0 ? P * (uint32_t)(nz - 1) : 0u;           // plane nz-1 or 0
#pragma unroll
          for (int r = 0; r < R; ++r) a.pz[r] = ld4(eo_prev + ra[r + 2] + pi);
        }
        if (PX) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const uint32_t rowb = ra[r + 2] - x0 + pk;
            a.pw[r] = eo_prev[rowb + (uint32_t)(nx - 1)];
            a.pe[r] = eo_prev[rowb];
          }
        }
      }
      const uint32_t o0 = poff(p), o1 = poff(p + 1);
      const bool v0 = pvalid(p), v1 = pvalid(p + 1);
#pragma unroll
      for (int i = 0; i < R + 2; ++i) cn[i] = e0at(i + 1, o1, v1);
      a.f[0] = e0at(0, o0, v0);
      a.f[1] = e0at(R + 3, o0, v0);
      // r32 at any cell where the sweep is not used (guard rows and planes)
      const uint32_t orr = v0 ? o0 : P * (uint32_t)ka;
#pragma unroll
      for (int i = 0; i < R + 2; ++i) a.r[i] = ld4(r32 + ra[i + 1] + orr);
    };
    // one Jacobi sweep of a row at one plane (k_mx_sweep's arithmetic):
    // centre c, z-neighbours zb/zt, y-neighbours n/s, right-hand side r; and,
    // when test, the stop-test residual of c there with the stale values
    // (sn, ss, st, sb when the flags are set; pw/pe across x)
    // one Jacobi sweep of a row at one plane (k_mx_sweep's arithmetic):
    // centre c, z-neighbours zb/zt, y-neighbours n/s, right-hand side r; and,
    // when test, the stop-test residual of c there.  zero: a flow guard row
    // or plane, whose correction is 0 (a wave-uniform branch).  Lanes past the
    // row end are not zeroed: their only reader is the east neighbour of the
    // lane holding x = nx - 1, which takes the guard (or the periodic image)
    // instead; padding cells are zeroed, as they are stored.
    auto sweep = [&](const float4 &c, const float4 &zb, const float4 &zt, const float4 &n,
                     const float4 &s, const float4 &r, bool zero, bool first1, bool test,
                     bool fN, const float4 &qn, bool fS, const float4 &qs, bool zT,
                     bool zB, const float4 &qz, float pw, float pe) -> float4 {
      if (zero) return zero4;
      const float ce[4] = {c.x, c.y, c.z, c.w};
      const float cn[4] = {n.x, n.y, n.z, n.w}, cs[4] = {s.x, s.y, s.z, s.w};
      const float ct[4] = {zt.x, zt.y, zt.z, zt.w}, cb[4] = {zb.x, zb.y, zb.z, zb.w};
      const float cr[4] = {r.x, r.y, r.z, r.w};
      const float wv = shr(c.w), ev = shl(c.x);
      float wrapW = 0.0f, wrapE = 0.0f;
      if (PX) {
        const float lastc = qw == 0 ? c.x : (qw == 1 ? c.y : (qw == 2 ? c.z : c.w));
        wrapW = bcast(lastc, Lw);
        wrapE = bcast(c.x, 0);
      }
      float ew[4], ee[4], v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ew[q] = q == 0 ? wv : ce[q - 1];
        if (hasf && q == 0) ew[q] = wrapW;
        if (has0 && q == 0) ew[q] = 0.0f;      // the lane below holds x = -1 (MULTI)
        ee[q] = q == 3 ? ev : ce[q + 1];
        // without padding, x = nx - 1 is the last cell of its lane
        if ((PAD || q == 3) && haslast && q == ql) ee[q] = PX ? wrapE : 0.0f;
        v[q] = jac_cell(cf, ee[q], ew[q], cn[q], cs[q], ct[q], cb[q], ce[q], cr[q]);
        if (first1) v[q] = -cf.af * cr[q];     // from e_0 = 0, as k_mx_sweep writes e_1
        if (PAD && !inrow(q)) v[q] = 0.0f;
      }
      if (test) {
        // stop-test residual; across the periodic boundaries the neighbours
        // are e_{j-1} (a wave-uniform branch: boundary rows and planes only,
        // or every row when x is periodic), elsewhere the sums are the update's
        float sq = 0.0f;
        if (PX || fN || fS || zT || zB) {
          const float qN[4] = {qn.x, qn.y, qn.z, qn.w}, qS[4] = {qs.x, qs.y, qs.z, qs.w};
          const float qZ[4] = {qz.x, qz.y, qz.z, qz.w};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float se = ((PAD || q == 3) && hasl && q == ql) ? pe : ee[q];
            const float sw = (hasf && q == 0) ? pw : ew[q];
            const float sn = fN ? qN[q] : cn[q];
            const float ss = fS ? qS[q] : cs[q];
            const float st_ = zT ? qZ[q] : ct[q];
            const float sb = zB ? qZ[q] : cb[q];
            const float e2 = 2.0f * ce[q];
            const float rr = cr[q] - ((se + sw - e2) * cf.idx2 + (sn + ss - e2) * cf.idy2 +
                                      (st_ + sb - e2) * cf.idz2);
            if (!PAD || q <= ql) sq = sq + rr * rr;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float e2 = 2.0f * ce[q];
            const float rr = cr[q] - ((ee[q] + ew[q] - e2) * cf.idx2 +
                                      (cn[q] + cs[q] - e2) * cf.idy2 +
                                      (ct[q] + cb[q] - e2) * cf.idz2);
            if (!PAD || q <= ql) sq = sq + rr * rr;
          }
        }
        if (!keep) sq = 0.0f;                  // the lane's 4 squares in FP32
        r2 = r2 + (double)sq;
      }
      return make_float4(v[0], v[1], v[2], v[3]);
    };
    // windows: e_j on rows y0-1..y0+R at planes p-1 (wm), p (wc) and p+1
    // (cn), four sets C0..C3 that rotate by name over four plane steps (the
    // fourth receives plane p+2 meanwhile); e_{j+1} on the wave's rows at
    // planes p-2 (u0) and p-1 (u1), on rows y0-1 and y0+R at plane p-1 (hs,
    // hn); r32 of the wave's rows at plane p-1 (rp)
    const int pa = ka - 1, pb = kb;                         // plane steps pa..pb
    float4 C0[R + 2], C1[R + 2], C2[R + 2], C3[R + 2], u0[R], u1[R], rp[R];
    float4 hs = zero4, hn = zero4;
    {
      const uint32_t oam = poff(pa - 1), oa = poff(pa);
      const bool vam = pvalid(pa - 1), va = pvalid(pa);
#pragma unroll
      for (int i = 0; i < R + 2; ++i) {
        C0[i] = e0at(i + 1, oam, vam);
        C1[i] = e0at(i + 1, oa, va);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        u0[r] = zero4;
        u1[r] = zero4;
        rp[r] = zero4;
      }
    }
    auto do_step = [&](int p, const StepIn &a, const float4 (&wm)[R + 2], const float4 (&wc)[R + 2],
                       const float4 (&cn)[R + 2]) {
      const bool pv = pvalid(p);
      const bool in = p >= ka && p < kb;
      const bool zT = sg.pz && p == nz - 1, zB = sg.pz && p == 0;
      // first sweep at plane p, rows y0-1..y0+R (index i: row y0-1+i)
      float4 v1[R + 2];
#pragma unroll
      for (int i = 0; i < R + 2; ++i) {
        const float4 &n = i + 1 < R + 2 ? wc[i + 1] : a.f[1];
        const float4 &s = i > 0 ? wc[i - 1] : a.f[0];
        if (i >= 1 && i <= R) {
          const int r = i - 1;
          v1[i] = sweep(wc[i], wm[i], cn[i], n, s, a.r[i], !pv || !rv[i + 1], first,
                        res && in && own[r], yN[r], a.pyn, yS[r], a.pys, zT, zB, a.pz[r],
                        a.pw[r], a.pe[r]);
          if (in && keep && own[r] && (edge[r] || zT || zB))
            st4(eo_out + ra[i + 1] + P * (uint32_t)p, v1[i]);
        } else {
          v1[i] = sweep(wc[i], wm[i], cn[i], n, s, a.r[i], !pv || !rv[i + 1], first, false,
                        false, zero4, false, zero4, false, false, zero4, 0.0f, 0.0f);
        }
      }
      // second sweep at plane p-1 on the wave's rows
      if (p - 1 >= ka) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float4 &n = r + 1 < R ? u1[r + 1] : hn;
          const float4 &s = r > 0 ? u1[r - 1] : hs;
          const float4 v2 = sweep(u1[r], u0[r], v1[r + 1], n, s, rp[r], false, false, false,
                                  false, zero4, false, zero4, false, false, zero4, 0.0f, 0.0f);
          if (keep && own[r]) st4(out + ra[r + 2] + P * (uint32_t)(p - 1), v2);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        u0[r] = u1[r];
        u1[r] = v1[r + 1];
        rp[r] = a.r[r + 1];
      }
      hs = v1[0];
      hn = v1[R + 1];
    };
    auto si = [&](StepIn &a, float4 (&cn)[R + 2], int p) { step_in(a, cn, p < pb ? p : pb); };
    StepIn A, B;
    A.pyn = A.pys = B.pyn = B.pys = zero4;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      A.pz[r] = B.pz[r] = zero4;
      A.pw[r] = A.pe[r] = B.pw[r] = B.pe[r] = 0.0f;
    }
    // step q reads (wm, wc, cn) = planes q-1, q, q+1; the loads of step q+1
    // go out before step q, into the set step q-1 last read
    si(A, C2, pa);
    for (int p = pa;; p += 4) {
      si(B, C3, p + 1);
      do_step(p, A, C0, C1, C2);
      if (p + 1 > pb) break;
      si(A, C0, p + 2);
      do_step(p + 1, B, C1, C2, C3);
      if (p + 2 > pb) break;
      si(B, C1, p + 3);
      do_step(p + 2, A, C2, C3, C0);
      if (p + 3 > pb) break;
      si(A, C2, p + 4);
      do_step(p + 3, B, C3, C0, C1);
      if (p + 4 > pb) break;
    }
  }
  const double sblk = block_sum(r2, sh);
  if (lin_tid() == 0 && res) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// One cell group of the two-sweep kernels with x not periodic: the arithmetic
// of k_mx_pair's sweep with no periodic image in x and rows of one segment.
// Centre c, z-neighbours zb/zt, y-neighbours n/s, right-hand side r; zero: a
// flow guard row or plane (or a row no reader needs), whose correction is 0;
// first1: e_0 = 0 (the sweep is -af r); test: add the stop-test residual of c
// to r2, with the stale values qn/qs/qz across periodic boundaries where the
// flags fN/fS/zT/zB are set.
template <bool PAD>
__device__ __forceinline__ float4 mx_cell4(const SweepCoef &cf, const float4 &c, const float4 &zb,
                                           const float4 &zt, const float4 &n, const float4 &s,
                                           const float4 &r, bool zero, bool first1, bool test,
                                           bool fN, const float *pn, bool fS, const float *ps,
                                           bool zT, bool zB, const float *pz, bool live,
                                           bool haslast, int ql, double &r2) {
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (zero) return zero4;
  const float ce[4] = {c.x, c.y, c.z, c.w};
  const float cn[4] = {n.x, n.y, n.z, n.w}, cs[4] = {s.x, s.y, s.z, s.w};
  const float ct[4] = {zt.x, zt.y, zt.z, zt.w}, cb[4] = {zb.x, zb.y, zb.z, zb.w};
  const float cr[4] = {r.x, r.y, r.z, r.w};
  const float wv = __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(c.w), 0x138, 0xF, 0xF, true));
  const float ev = __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(c.x), 0x130, 0xF, 0xF, true));
  float ew[4], ee[4], v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ew[q] = q == 0 ? wv : ce[q - 1];
    ee[q] = q == 3 ? ev : ce[q + 1];
    if ((PAD || q == 3) && haslast && q == ql) ee[q] = 0.0f;
    v[q] = jac_cell(cf, ee[q], ew[q], cn[q], cs[q], ct[q], cb[q], ce[q], cr[q]);
    if (first1) v[q] = -cf.af * cr[q];
    if (PAD && !(live && q <= ql)) v[q] = 0.0f;
  }
  if (test) {
    float sq = 0.0f;
    if (fN || fS || zT || zB) {
      const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const float4 qn = fN ? ld4(pn) : zero4, qs = fS ? ld4(ps) : zero4;
      const float4 qz = (zT || zB) ? ld4(pz) : zero4;
      const float qN[4] = {qn.x, qn.y, qn.z, qn.w}, qS[4] = {qs.x, qs.y, qs.z, qs.w};
      const float qZ[4] = {qz.x, qz.y, qz.z, qz.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float sn = fN ? qN[q] : cn[q];
        const float ss = fS ? qS[q] : cs[q];
        const float st_ = zT ? qZ[q] : ct[q];
        const float sb = zB ? qZ[q] : cb[q];
        const float e2 = 2.0f * ce[q];
        const float rr = cr[q] - ((ee[q] + ew[q] - e2) * cf.idx2 + (sn + ss - e2) * cf.idy2 +
                                  (st_ + sb - e2) * cf.idz2);
        if (!PAD || q <= ql) sq = sq + rr * rr;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float e2 = 2.0f * ce[q];
        const float rr = cr[q] - ((ee[q] + ew[q] - e2) * cf.idx2 + (cn[q] + cs[q] - e2) * cf.idy2 +
                                  (ct[q] + cb[q] - e2) * cf.idz2);
        if (!PAD || q <= ql) sq = sq + rr * rr;
      }
    }
    if (!live) sq = 0.0f;
    r2 = r2 + (double)sq;
  }
  return make_float4(v[0], v[1], v[2], v[3]);
}

// The two-sweep kernel with rows shared through LDS, for x not periodic and
// rows of at most one segment (nx <= 256).  A block of kLdsWaves waves owns a
// tile of at most kLdsRows consecutive rows ya..yb-1 and marches a z chunk.
// Wave w computes the first sweep, e_{j+1}, on rows ya-1+2w and ya+2w (the
// tile's rows and one more on each side), and the second sweep, e_{j+2}, on
// those of them that are the tile's.  Each wave loads e_j and r32 of its two
// rows only (wave 0 and the last wave also e_j of the row beyond the first or
// last computed row), keeps them in registers for the centre and the z
// neighbours, and hands them to the neighbouring waves through LDS: e_j of
// plane p+1 and e_{j+1} of plane p at plane step p, one barrier per step.
// Against k_mx_pair (each wave recomputes the rows next to its own and loads
// them again), the first sweep runs on 16 rows for 14 and every row of e_j
// and r32 is loaded once per tile, and a wave holds at most 128 VGPRs, so
// that 16 waves per CU stay resident to cover HBM latency (k_mx_pair: 241,
// 8 waves).  Every cell's arithmetic is k_mx_sweep's, so e_{j+2} is bit for
// bit what two launches of it give, and the boundary copies of e_{j+1} are
// those of k_mx_pair.  The host sends the launches it marks for the stop test
// to k_mx_pair; this kernel evaluates the stop-test residual of e_j (cell for
// cell k_mx_pair's) only while the stop is near, in a pass of its own ahead of
// the sweeps, which keeps the sweeps' registers within 128.
constexpr int kLdsWaves = 8;                  // waves per block
constexpr int kLdsDepth = 1;                  // plane steps the loads run ahead
constexpr int kLdsRows = 2 * kLdsWaves - 2;   // rows of a tile

// NW waves per block (2 NW rows of the first sweep, 2 NW - 2 of a tile),
// loads issued D plane steps ahead; at most 128 VGPRs for D = 1, so that
// two blocks of 8 waves are resident per CU.  At 256^3: 8 waves, D = 1:
// 36.3 us per launch (median); 6 waves 38.0; D = 2 (145 VGPRs, 3 waves per
// SIMD) 46 us with 8 waves, 50 us with 6, as fewer blocks stay resident.

template <bool PAD, bool FIRST, int NW, int D>
__global__ void __launch_bounds__(64 * NW, D == 1 ? 4 : 3) k_mx_pair_lds(
    SweepGeo sg, SweepCoef cf, const float *__restrict__ r32, const float *__restrict__ e,
    float *__restrict__ out, float *__restrict__ eo_out, const float *__restrict__ eo_prev,
    const double *__restrict__ Pin, double *__restrict__ Pout, JacobiState *st, MxLaunch L,
    unsigned long long *ks) {
  constexpr int kLdsCRows = 2 * NW;
  __shared__ double sh[NW];
  __shared__ int decide;
  // e_j of plane q in EJs[q & 1] (row i of the tile at index i + 1, i = -1..kLdsCRows);
  // e_{j+1} of plane q in E1s[q & 1]
  __shared__ float4 EJs[2][kLdsCRows + 2][64];
  __shared__ float4 E1s[2][kLdsCRows][64];
  kstamp_begin(ks);
  StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (!cur.done && (int64_t)L.base + L.j >= st->last) {   // iterate 2*maxit+1 reached (:58)
    cur.done = 1;
    cur.result = st->last;
    cur.buf = L.ibuf;
  }
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }
  const bool res = !FIRST && (L.test || cur.near);
  const uint32_t NX = sg.NX, P = sg.P;
  const int nx = sg.nx, ny = sg.ny, nz = sg.nz;
  const int lane = (int)threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t zrow = P * (uint32_t)nz;                   // the zero row of E
  const int xs = 4 * lane;
  const bool live = xs < nx;
  const uint32_t x0 = live ? (uint32_t)xs : 0u;
  const uint32_t zr = zrow + x0;
  const int ql = nx - 1 - xs;
  const bool haslast = live && ql < 4;
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  auto wrapz = [&](int pp) {
    int kk = pp < 0 ? pp + nz : pp;
    kk = kk < 0 ? kk + nz : kk;
    kk = kk >= nz ? kk - nz : kk;
    return kk >= nz ? kk - nz : kk;
  };
  auto pvalid = [&](int pp) { return sg.pz | ((pp >= 0) & (pp < nz)); };
  auto poff = [&](int pp) { return P * (uint32_t)wrapz(pp); };
  const int i0 = 2 * wv, i1 = 2 * wv + 1;                   // the wave's rows of the tile
  const int ix = wv == 0 ? -1 : kLdsCRows;                  // its extra row of e_j
  for (int t = w; t < sg.TP; t += sg.W) {
    // tile t: z chunk zi of the XCD slab, rows ya..yb-1 (block-uniform)
    const int zi = t % sg.nzc, ty = t / sg.nzc;
    const int ka = z0 + ((z1 - z0) * zi) / sg.nzc, kb = z0 + ((z1 - z0) * (zi + 1)) / sg.nzc;
    const int ya = (ty * ny) / sg.nty, yb = ((ty + 1) * ny) / sg.nty;
    const int Yo = yb - ya;
    if (Yo <= 0 || kb <= ka) continue;
    // row i of the tile is y = ya - 1 + i: its address (the periodic image,
    // or row ya for a flow guard row) and whether it is in the grid
    auto row_at = [&](int i, bool &ok) -> uint32_t {
      int yy = ya - 1 + i;
      ok = true;
      if (yy < 0 || yy >= ny) {
        if (sg.py) {
          yy = yy % ny;
          if (yy < 0) yy += ny;
        } else {
          ok = false;
          yy = ya;
        }
      }
      return x0 + NX * (uint32_t)yy;
    };
    bool ok0, ok1, okx;
    const uint32_t ra0 = row_at(i0, ok0), ra1 = row_at(i1, ok1), rax = row_at(ix, okx);
    // e_j is read at rows -1..Yo+2, the first sweep computed at rows 0..Yo+1
    // (zero at flow guard rows and past Yo+1), the second at rows 1..Yo
    const bool rok0 = !FIRST && live && ok0 && i0 <= Yo + 2;
    const bool rok1 = !FIRST && live && ok1 && i1 <= Yo + 2;
    const bool hasx = !FIRST && (wv == 0 || (wv == NW - 1 && Yo + 2 >= kLdsCRows));
    const bool rokx = hasx && live && okx;
    const bool cp0 = ok0 && i0 <= Yo + 1, cp1 = ok1 && i1 <= Yo + 1;
    const bool op0 = i0 >= 1 && i0 <= Yo, op1 = i1 >= 1 && i1 <= Yo;
    const int y0r = ya - 1 + i0, y1r = ya - 1 + i1;
    const bool yN0 = sg.py && y0r == ny - 1, yS0 = sg.py && y0r == 0;
    const bool yN1 = sg.py && y1r == ny - 1, yS1 = sg.py && y1r == 0;
    auto eld = [&](uint32_t ra, bool rok, int pp) -> float4 {
      if constexpr (FIRST) return zero4;      // e_0 = 0: nothing to read
      return ld4(e + ((pvalid(pp) & rok) ? ra + poff(pp) : zr));
    };
    // the stop-test residual of e_j on the tile's rows (launches while the
    // stop is near), in a pass of its own ahead of the sweeps, so that the
    // sweeps hold no registers for it: across periodic boundaries the
    // neighbours are e_{j-1} (rows 0 and ny-1, planes 0 and nz-1)
    if (res) {
      // row by row, the loads of plane p+1 issued before the residual of
      // plane p
#pragma unroll 1
      for (int r = 0; r < 2; ++r) {
        if (!(r ? op1 : op0)) continue;                      // wave-uniform
        const int i = r ? i1 : i0;
        const uint32_t rc = r ? ra1 : ra0;
        bool okn, oks;
        const uint32_t rn = row_at(i + 1, okn), rs = row_at(i - 1, oks);
        const bool fN = r ? yN1 : yN0, fS = r ? yS1 : yS0;
        struct RIn {
          float4 t, n, s, q;
        };
        auto rissue = [&](RIn &a, int p) {
          a.t = eld(rc, live, p + 1);
          a.n = eld(rn, live && okn, p);
          a.s = eld(rs, live && oks, p);
          a.q = ld4(r32 + rc + poff(p));
        };
        float4 b = eld(rc, live, ka - 1), cc = eld(rc, live, ka);
        auto rstep = [&](int p, const RIn &a) {
          const uint32_t pk = P * (uint32_t)p;
          const uint32_t pi = p == 0 ? P * (uint32_t)(nz - 1) : 0u;    // plane nz-1 or 0
          (void)mx_cell4<PAD>(cf, cc, b, a.t, a.n, a.s, a.q, false, false, true, fN,
                              eo_prev + x0 + pk, fS, eo_prev + x0 + NX * (uint32_t)(ny - 1) + pk,
                              sg.pz && p == nz - 1, sg.pz && p == 0, eo_prev + rc + pi, live,
                              haslast, ql, r2);
          b = cc;
          cc = a.t;
        };
        RIn RA, RB;
        rissue(RA, ka);
        for (int p = ka;; p += 2) {
          rissue(RB, p + 1 < kb ? p + 1 : kb - 1);
          rstep(p, RA);
          if (p + 1 >= kb) break;
          rissue(RA, p + 2 < kb ? p + 2 : kb - 1);
          rstep(p + 1, RB);
          if (p + 2 >= kb) break;
        }
      }
    }
    // the march: step k = 0..K has its front (first sweep) at plane f(k),
    // up from ka-1 to kb in even chunks, down from kb to ka-1 in odd ones, so
    // that neighbouring chunks of the slab read their shared planes at the
    // same time (both at the start of their march or both at its end) and
    // the second reads them from L2.  Going down swaps only the z-neighbours
    // of each cell, whose sum (t + b) does not depend on their order.
    const bool up = (zi & 1) == 0;
    const int dz = up ? 1 : -1, K = kb - ka + 1;
    auto front = [&](int k) { return up ? ka - 1 + k : kb - k; };
    // registers: e_j of the wave's rows at f(k) (c*), e_{j+1} at f(k-1)
    // (u1*), r32 at f(k-1) (rp*).  e_j at f(k-1) and e_{j+1} at f(k-2) of the
    // wave's rows are read back from LDS, where only this wave writes them
    // (before it overwrites them with f(k+1) and f(k)).  EJs[k & 1] holds e_j
    // at f(k), E1s[k & 1] e_{j+1} at f(k).
    float4 c0 = eld(ra0, rok0, front(0)), c1 = eld(ra1, rok1, front(0));
    float4 u10 = zero4, u11 = zero4, rp0 = zero4, rp1 = zero4;
    if (!FIRST) {
      EJs[1][i0 + 1][lane] = eld(ra0, rok0, front(0) - dz);
      EJs[1][i1 + 1][lane] = eld(ra1, rok1, front(0) - dz);
      EJs[0][i0 + 1][lane] = c0;
      EJs[0][i1 + 1][lane] = c1;
      if (hasx) EJs[0][ix + 1][lane] = eld(rax, rokx, front(0));
    }
    // loads for step k: e_j of the wave's rows (and its extra row) at
    // f(k+1), r32 of its rows at f(k) (plane ka where no sweep uses it)
    struct In {
      float4 e0, e1, ex, r0, r1;
    };
    auto issue = [&](In &a, int k) {
      const int f = front(k);
      a.e0 = eld(ra0, rok0, f + dz);
      a.e1 = eld(ra1, rok1, f + dz);
      if (hasx) a.ex = eld(rax, rokx, f + dz);
      const uint32_t orr = pvalid(f) ? poff(f) : P * (uint32_t)ka;
      a.r0 = ld4(r32 + ra0 + orr);
      a.r1 = ld4(r32 + ra1 + orr);
    };
    In S[D + 1];
#pragma unroll
    for (int q = 0; q <= D; ++q) S[q].ex = zero4;
#pragma unroll
    for (int q = 0; q < D; ++q) issue(S[q], q < K ? q : K);
    __syncthreads();
    auto step = [&](int k, const In &a) {
      const int p = front(k);
      const bool pv = pvalid(p);
      const bool in = p >= ka && p < kb;
      const bool zT = sg.pz && p == nz - 1, zB = sg.pz && p == 0;
      const int s = k & 1;
      // e_j at f(k) of the rows next to the wave's (written by its
      // neighbours), and of its own rows at f(k-1)
      const float4 s0 = FIRST ? zero4 : EJs[s][i0][lane];
      const float4 n1 = FIRST ? zero4 : EJs[s][i1 + 2][lane];
      const float4 m0 = FIRST ? zero4 : EJs[s ^ 1][i0 + 1][lane];
      const float4 m1 = FIRST ? zero4 : EJs[s ^ 1][i1 + 1][lane];
      // first sweep at f(k)
      const float4 v0 = mx_cell4<PAD>(cf, c0, m0, a.e0, c1, s0, a.r0, !pv || !cp0, FIRST, false,
                                      false, nullptr, false, nullptr, false, false, nullptr, live,
                                      haslast, ql, r2);
      const float4 v1 = mx_cell4<PAD>(cf, c1, m1, a.e1, n1, c0, a.r1, !pv || !cp1, FIRST, false,
                                      false, nullptr, false, nullptr, false, false, nullptr, live,
                                      haslast, ql, r2);
      if (in && live) {
        if (op0 && (yN0 || yS0 || zT || zB)) st4(eo_out + ra0 + P * (uint32_t)p, v0);
        if (op1 && (yN1 || yS1 || zT || zB)) st4(eo_out + ra1 + P * (uint32_t)p, v1);
      }
      // second sweep at f(k-1) on the tile's rows
      if (k >= 2) {
        const uint32_t po = P * (uint32_t)(p - dz);
        if (op0) {
          const float4 u00 = E1s[s][i0][lane];                 // f(k-2)
          const float4 sv = E1s[s ^ 1][i0 - 1][lane];
          const float4 v = mx_cell4<PAD>(cf, u10, u00, v0, u11, sv, rp0, false, false, false,
                                         false, nullptr, false, nullptr, false, false, nullptr,
                                         live, haslast, ql, r2);
          if (live) st4(out + ra0 + po, v);
        }
        if (op1) {
          const float4 u01 = E1s[s][i1][lane];                 // f(k-2)
          const float4 n = E1s[s ^ 1][i1 + 1][lane];
          const float4 v = mx_cell4<PAD>(cf, u11, u01, v1, n, u10, rp1, false, false, false,
                                         false, nullptr, false, nullptr, false, false, nullptr,
                                         live, haslast, ql, r2);
          if (live) st4(out + ra1 + po, v);
        }
      }
      // hand on e_j at f(k+1) and e_{j+1} at f(k)
      if (!FIRST) {
        EJs[s ^ 1][i0 + 1][lane] = a.e0;
        EJs[s ^ 1][i1 + 1][lane] = a.e1;
        if (hasx) EJs[s ^ 1][ix + 1][lane] = a.ex;
      }
      E1s[s][i0][lane] = v0;
      E1s[s][i1][lane] = v1;
      u10 = v0;
      u11 = v1;
      rp0 = a.r0;
      rp1 = a.r1;
      c0 = a.e0;
      c1 = a.e1;
      __syncthreads();
    };
    // the loads of step k+D go out before step k (into the set step k-1 used)
    bool more = true;
    for (int k0 = 0; more; k0 += D + 1) {
#pragma unroll
      for (int u = 0; u <= D; ++u) {
        const int k = k0 + u;
        if (more) {
          issue(S[(u + D) % (D + 1)], k + D < K ? k + D : K);
          step(k, S[u]);
          more = k + 1 <= K;
        }
      }
    }
  }
  const double sblk = block_sum(r2, sh);
  if (lin_tid() == 0 && res) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// The three-sweep kernel: k_mx_pair_lds's scheme one level deeper, e_j ->
// e_{j+3} with the HBM traffic of one sweep (read e_j and r32, write
// e_{j+3}).  A block of NW waves owns a tile of at most 2 NW - 4 rows
// ya..yb-1; wave w computes the first sweep on rows ya-2+2w and ya-1+2w at
// the march front f(k), the second sweep on those of them in ya-1..yb at
// f(k-1), the third on those in ya..yb-1 at f(k-2), and hands each level's
// rows to its neighbours through LDS, one barrier per plane step.  The z
// chunk's march covers planes ka-2..kb+1 for the first sweep.  Every cell's
// arithmetic is k_mx_sweep's (e_{j+3} is bit for bit what three single
// sweeps give); the boundary copies it leaves are those of e_{j+2}, which the
// next launch's stop test (of e_{j+3}) takes as the stale halo.  It does not
// evaluate the stop test except while the stop is near (the host sends the
// launches it marks for the test to k_mx_pair), in a pass of its own as
// k_mx_pair_lds.  With 6 waves per block, 76 KB of LDS and at most 168 VGPRs,
// two blocks are resident per CU.  Measured at 256^3: 68 us per launch, 22.6
// us per sweep against 18.9 with two sweeps per launch (k_mx_pair_lds: 36 us
// per launch): the three levels' chain per plane step, the LDS traffic
// (26 accesses of 1 KB per wave and step, about 1 us of a CU's LDS bandwidth)
// and the tiles of 8 rows (14 rows of e_j read per 8) cost more than the
// bytes saved, so it runs only when asked for (picard_split_sweeps_per_launch
// with 3).
constexpr int kTriWaves = 6;                   // waves per block
constexpr int kTriRows = 2 * kTriWaves - 4;    // rows of a tile

template <bool PAD, bool FIRST, int NW>
__global__ void __launch_bounds__(64 * NW, 3) k_mx_triple_lds(
    SweepGeo sg, SweepCoef cf, const float *__restrict__ r32, const float *__restrict__ e,
    float *__restrict__ out, float *__restrict__ eo_out, const float *__restrict__ eo_prev,
    const double *__restrict__ Pin, double *__restrict__ Pout, JacobiState *st, MxLaunch L,
    unsigned long long *ks) {
  constexpr int CR = 2 * NW;                    // rows of the first sweep
  __shared__ double sh[NW];
  __shared__ int decide;
  // e_j at f(k) in EJs[k & 1] (row i of the tile at index i + 1, i = -1..CR);
  // e_{j+1} at f(k) in E1s[k & 1]; e_{j+2} at f(k-1) in E2s[k & 1]
  __shared__ float4 EJs[2][CR + 2][64];
  __shared__ float4 E1s[2][CR][64];
  __shared__ float4 E2s[2][CR][64];
  kstamp_begin(ks);
  StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (!cur.done && (int64_t)L.base + L.j >= st->last) {   // iterate 2*maxit+1 reached (:58)
    cur.done = 1;
    cur.result = st->last;
    cur.buf = L.ibuf;
  }
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }
  const bool res = !FIRST && (L.test || cur.near);
  const uint32_t NX = sg.NX, P = sg.P;
  const int nx = sg.nx, ny = sg.ny, nz = sg.nz;
  const int lane = (int)threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t zrow = P * (uint32_t)nz;                   // the zero row of E
  const int xs = 4 * lane;
  const bool live = xs < nx;
  const uint32_t x0 = live ? (uint32_t)xs : 0u;
  const uint32_t zr = zrow + x0;
  const int ql = nx - 1 - xs;
  const bool haslast = live && ql < 4;
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  auto wrapz = [&](int pp) {
    int kk = pp < 0 ? pp + nz : pp;
    kk = kk < 0 ? kk + nz : kk;
    kk = kk < 0 ? kk + nz : kk;
    kk = kk >= nz ? kk - nz : kk;
    kk = kk >= nz ? kk - nz : kk;
    return kk >= nz ? kk - nz : kk;
  };
  auto pvalid = [&](int pp) { return sg.pz | ((pp >= 0) & (pp < nz)); };
  auto poff = [&](int pp) { return P * (uint32_t)wrapz(pp); };
  const int i0 = 2 * wv, i1 = 2 * wv + 1;                   // the wave's rows of the tile
  const int ix = wv == 0 ? -1 : CR;                         // its extra row of e_j
  for (int t = w; t < sg.TP; t += sg.W) {
    const int zi = t % sg.nzc, ty = t / sg.nzc;
    const int ka = z0 + ((z1 - z0) * zi) / sg.nzc, kb = z0 + ((z1 - z0) * (zi + 1)) / sg.nzc;
    const int ya = (ty * ny) / sg.nty, yb = ((ty + 1) * ny) / sg.nty;
    const int Yo = yb - ya;
    if (Yo <= 0 || kb <= ka) continue;
    // row i of the tile is y = ya - 2 + i
    auto row_at = [&](int i, bool &ok) -> uint32_t {
      int yy = ya - 2 + i;
      ok = true;
      if (yy < 0 || yy >= ny) {
        if (sg.py) {
          yy = yy % ny;
          if (yy < 0) yy += ny;
        } else {
          ok = false;
          yy = ya;
        }
      }
      return x0 + NX * (uint32_t)yy;
    };
    bool ok0, ok1, okx;
    const uint32_t ra0 = row_at(i0, ok0), ra1 = row_at(i1, ok1), rax = row_at(ix, okx);
    // e_j is read at rows -1..Yo+4; the first sweep at rows 0..Yo+3, the
    // second at 1..Yo+2, the third (the tile's rows) at 2..Yo+1
    const bool rok0 = !FIRST && live && ok0 && i0 <= Yo + 4;
    const bool rok1 = !FIRST && live && ok1 && i1 <= Yo + 4;
    const bool hasx = !FIRST && (wv == 0 || (wv == NW - 1 && Yo + 4 >= CR));
    const bool rokx = hasx && live && okx;
    const bool a0 = ok0 && i0 <= Yo + 3, a1 = ok1 && i1 <= Yo + 3;
    const bool b0 = ok0 && i0 >= 1 && i0 <= Yo + 2, b1 = ok1 && i1 >= 1 && i1 <= Yo + 2;
    const bool op0 = i0 >= 2 && i0 <= Yo + 1, op1 = i1 >= 2 && i1 <= Yo + 1;
    const int y0r = ya - 2 + i0, y1r = ya - 2 + i1;
    const bool yN0 = sg.py && y0r == ny - 1, yS0 = sg.py && y0r == 0;
    const bool yN1 = sg.py && y1r == ny - 1, yS1 = sg.py && y1r == 0;
    auto eld = [&](uint32_t ra, bool rok, int pp) -> float4 {
      if constexpr (FIRST) return zero4;      // e_0 = 0: nothing to read
      return ld4(e + ((pvalid(pp) & rok) ? ra + poff(pp) : zr));
    };
    if (res) {
      // the stop-test residual of e_j on the tile's rows (launches while the
      // stop is near), in a pass of its own, row by row, the loads of plane
      // p+1 issued before the residual of plane p
#pragma unroll 1
      for (int r = 0; r < 2; ++r) {
        if (!(r ? op1 : op0)) continue;                      // wave-uniform
        const int i = r ? i1 : i0;
        const uint32_t rc = r ? ra1 : ra0;
        bool okn, oks;
        const uint32_t rn = row_at(i + 1, okn), rs = row_at(i - 1, oks);
        const bool fN = r ? yN1 : yN0, fS = r ? yS1 : yS0;
        struct RIn {
          float4 t, n, s, q;
        };
        auto rissue = [&](RIn &a, int p) {
          a.t = eld(rc, live, p + 1);
          a.n = eld(rn, live && okn, p);
          a.s = eld(rs, live && oks, p);
          a.q = ld4(r32 + rc + poff(p));
        };
        float4 b = eld(rc, live, ka - 1), cc = eld(rc, live, ka);
        auto rstep = [&](int p, const RIn &a) {
          const uint32_t pk = P * (uint32_t)p;
          const uint32_t pi = p == 0 ? P * (uint32_t)(nz - 1) : 0u;    // plane nz-1 or 0
          (void)mx_cell4<PAD>(cf, cc, b, a.t, a.n, a.s, a.q, false, false, true, fN,
                              eo_prev + x0 + pk, fS, eo_prev + x0 + NX * (uint32_t)(ny - 1) + pk,
                              sg.pz && p == nz - 1, sg.pz && p == 0, eo_prev + rc + pi, live,
                              haslast, ql, r2);
          b = cc;
          cc = a.t;
        };
        RIn RA, RB;
        rissue(RA, ka);
        for (int p = ka;; p += 2) {
          rissue(RB, p + 1 < kb ? p + 1 : kb - 1);
          rstep(p, RA);
          if (p + 1 >= kb) break;
          rissue(RA, p + 2 < kb ? p + 2 : kb - 1);
          rstep(p + 1, RB);
          if (p + 2 >= kb) break;
        }
      }
    }
    // the march: fronts f(0..K), up from ka-2 to kb+1 in even chunks, down
    // in odd ones
    const bool up = (zi & 1) == 0;
    const int dz = up ? 1 : -1, K = kb - ka + 3;
    auto front = [&](int k) { return up ? ka - 2 + k : kb + 1 - k; };
    // registers: e_j of the wave's rows at f(k) (c*), e_{j+1} at f(k-1)
    // (u*), e_{j+2} at f(k-2) (x*), r32 at f(k-1) (rp*) and f(k-2) (rq*);
    // the level below each one's centre plane comes back from LDS, where
    // only this wave writes its rows
    float4 c0 = eld(ra0, rok0, front(0)), c1 = eld(ra1, rok1, front(0));
    float4 u0 = zero4, u1 = zero4, x0v = zero4, x1v = zero4;
    float4 rp0 = zero4, rp1 = zero4, rq0 = zero4, rq1 = zero4;
    if (!FIRST) {
      EJs[1][i0 + 1][lane] = eld(ra0, rok0, front(0) - dz);
      EJs[1][i1 + 1][lane] = eld(ra1, rok1, front(0) - dz);
      EJs[0][i0 + 1][lane] = c0;
      EJs[0][i1 + 1][lane] = c1;
      if (hasx) EJs[0][ix + 1][lane] = eld(rax, rokx, front(0));
    }
    struct In {
      float4 e0, e1, ex, r0, r1;
    };
    auto issue = [&](In &a, int k) {
      const int f = front(k);
      a.e0 = eld(ra0, rok0, f + dz);
      a.e1 = eld(ra1, rok1, f + dz);
      if (hasx) a.ex = eld(rax, rokx, f + dz);
      const uint32_t orr = pvalid(f) ? poff(f) : P * (uint32_t)ka;
      a.r0 = ld4(r32 + ra0 + orr);
      a.r1 = ld4(r32 + ra1 + orr);
    };
    In A, B;
    A.ex = B.ex = zero4;
    issue(A, 0);
    __syncthreads();
    auto step = [&](int k, const In &a) {
      const int p = front(k);
      const int s = k & 1;
      // first sweep at f(k)
      const bool pv = pvalid(p);
      const float4 s0 = FIRST ? zero4 : EJs[s][i0][lane];
      const float4 n1 = FIRST ? zero4 : EJs[s][i1 + 2][lane];
      const float4 m0 = FIRST ? zero4 : EJs[s ^ 1][i0 + 1][lane];
      const float4 m1 = FIRST ? zero4 : EJs[s ^ 1][i1 + 1][lane];
      const float4 v0 = mx_cell4<PAD>(cf, c0, m0, a.e0, c1, s0, a.r0, !pv || !a0, FIRST, false,
                                      false, nullptr, false, nullptr, false, false, nullptr, live,
                                      haslast, ql, r2);
      const float4 v1 = mx_cell4<PAD>(cf, c1, m1, a.e1, n1, c0, a.r1, !pv || !a1, FIRST, false,
                                      false, nullptr, false, nullptr, false, false, nullptr, live,
                                      haslast, ql, r2);
      // second sweep at f(k-1)
      float4 w0 = zero4, w1 = zero4;
      if (k >= 2) {
        const int q = p - dz;
        const bool qv = pvalid(q);
        const float4 ub0 = E1s[s][i0][lane], ub1 = E1s[s][i1][lane];        // f(k-2)
        const float4 us0 = wv > 0 ? E1s[s ^ 1][i0 - 1][lane] : zero4;
        const float4 un1 = wv < NW - 1 ? E1s[s ^ 1][i1 + 1][lane] : zero4;
        w0 = mx_cell4<PAD>(cf, u0, ub0, v0, u1, us0, rp0, !qv || !b0, false, false, false,
                           nullptr, false, nullptr, false, false, nullptr, live, haslast, ql, r2);
        w1 = mx_cell4<PAD>(cf, u1, ub1, v1, un1, u0, rp1, !qv || !b1, false, false, false,
                           nullptr, false, nullptr, false, false, nullptr, live, haslast, ql, r2);
        if (live && q >= ka && q < kb) {
          const bool zT = sg.pz && q == nz - 1, zB = sg.pz && q == 0;
          if (op0 && (yN0 || yS0 || zT || zB)) st4(eo_out + ra0 + P * (uint32_t)q, w0);
          if (op1 && (yN1 || yS1 || zT || zB)) st4(eo_out + ra1 + P * (uint32_t)q, w1);
        }
      }
      // third sweep at f(k-2) on the tile's rows
      if (k >= 4) {
        const uint32_t po = P * (uint32_t)(p - 2 * dz);
        const float4 xb0 = E2s[s][i0][lane], xb1 = E2s[s][i1][lane];        // f(k-3)
        if (op0) {
          const float4 xs0 = E2s[s ^ 1][i0 - 1][lane];
          const float4 y = mx_cell4<PAD>(cf, x0v, xb0, w0, x1v, xs0, rq0, false, false, false,
                                         false, nullptr, false, nullptr, false, false, nullptr,
                                         live, haslast, ql, r2);
          if (live) st4(out + ra0 + po, y);
        }
        if (op1) {
          const float4 xn1 = E2s[s ^ 1][i1 + 1][lane];
          const float4 y = mx_cell4<PAD>(cf, x1v, xb1, w1, xn1, x0v, rq1, false, false, false,
                                         false, nullptr, false, nullptr, false, false, nullptr,
                                         live, haslast, ql, r2);
          if (live) st4(out + ra1 + po, y);
        }
      }
      // hand on e_j at f(k+1), e_{j+1} at f(k), e_{j+2} at f(k-1)
      if (!FIRST) {
        EJs[s ^ 1][i0 + 1][lane] = a.e0;
        EJs[s ^ 1][i1 + 1][lane] = a.e1;
        if (hasx) EJs[s ^ 1][ix + 1][lane] = a.ex;
      }
      E1s[s][i0][lane] = v0;
      E1s[s][i1][lane] = v1;
      E2s[s][i0][lane] = w0;
      E2s[s][i1][lane] = w1;
      x0v = w0;
      x1v = w1;
      u0 = v0;
      u1 = v1;
      rq0 = rp0;
      rq1 = rp1;
      rp0 = a.r0;
      rp1 = a.r1;
      c0 = a.e0;
      c1 = a.e1;
      __syncthreads();
    };
    for (int k = 0;; k += 2) {
      issue(B, k + 1 < K ? k + 1 : K);
      step(k, A);
      if (k + 1 > K) break;
      issue(A, k + 2 < K ? k + 2 : K);
      step(k + 1, B);
      if (k + 2 > K) break;
    }
  }
  const double sblk = block_sum(r2, sh);
  if (lin_tid() == 0 && res) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// End of an outer step: the stop test of its last inner iterate and, when
// the solve ended inside this step (converged, or the last iterate reached),
// u += e_k for that iterate's k; otherwise nothing (the next outer step
// applies e_m).
__global__ void __launch_bounds__(kStencilT) k_mx_update(GridDims g, CompactDims cd, TileSched ts,
                                                    double *__restrict__ u, const float *E0,
                                                    const float *E1, const double *__restrict__ Pin,
                                                    JacobiState *st, MxLaunch L) {
  __shared__ double sh[kStencilT / 64];
  __shared__ int decide;
  const StopSlot prev = st->slot[(L.seq - 1) & 1];
  StopSlot cur = mx_decide(st, L, Pin, sh, &decide);
  if (!cur.done && L.base + L.m >= st->last) {
    cur.done = 1;
    cur.result = st->last;
    cur.buf = L.ibuf;
  }
  if (blockIdx.x == 0 && lin_tid() == 0) st->slot[L.seq & 1] = cur;
  // solves that ended before this outer step have nothing left to apply
  if (prev.done && prev.result < L.base) return;
  if (!cur.done) return;
  const int64_t k = cur.result - L.base;
  if (k <= 0) return;
  const float *e = cur.buf ? E1 : E0;
  for_my_cells(ts, g.nx, g.ny, g.nz, [&](int i, int j, int kk, bool valid) {
    if (!valid) return;
    const int64_t c = node_index(g, i, j, kk);
    u[c] = u[c] + (double)e[compact_index(cd, i, j, kk)];
  });
}

// Periodic halo fill of U in place: every guard node of a periodic dimension
// takes the value at its composite image (z, then y, then x plane copies).
__global__ void __launch_bounds__(kT) k_mx_halo(GridDims g, double *__restrict__ U) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int si = i, sj = j, sk = k;
  if (!g.flow[0]) si = i == 1 ? g.nx + 1 : (i == g.nx + 2 ? 2 : i);
  if (!g.flow[1]) sj = j == 1 ? g.ny + 1 : (j == g.ny + 2 ? 2 : j);
  if (!g.flow[2]) sk = k == 1 ? g.nz + 1 : (k == g.nz + 2 ? 2 : k);
  const int64_t src = node_index(g, si, sj, sk);
  if (src != node) U[node] = U[src];
}

// E = -grad U as FP32, with the same halo rules as the FP64 form.
__global__ void __launch_bounds__(kT) k_sp_efield(GridDims g, const double *__restrict__ U,
                                                   float *__restrict__ E) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int si = i, sj = j, sk = k;
  if (!g.flow[0]) si = i == 1 ? g.nx + 1 : (i == g.nx + 2 ? 2 : i);
  if (!g.flow[1]) sj = j == 1 ? g.ny + 1 : (j == g.ny + 2 ? 2 : j);
  if (!g.flow[2]) sk = k == 1 ? g.nz + 1 : (k == g.nz + 2 ? 2 : k);
  double ex = 0.0, ey = 0.0, ez = 0.0;
  if (si >= 2 && si <= g.nx + 1 && sj >= 2 && sj <= g.ny + 1 && sk >= 2 && sk <= g.nz + 1) {
    const int64_t c = node_index(g, si, sj, sk);
    const int64_t X = g.X, XY = (int64_t)g.X * g.Y;
    ez = -(U[c + XY] - U[c - XY]) * 0.5 / g.dxyz[2];
    ey = -(U[c + X] - U[c - X]) * 0.5 / g.dxyz[1];
    ex = -(U[c + 1] - U[c - 1]) * 0.5 / g.dxyz[0];
  }
  E[4 * node + 0] = (float)ex;
  E[4 * node + 1] = (float)ey;
  E[4 * node + 2] = (float)ez;
  E[4 * node + 3] = 0.0f;
}

struct BorisF32 {
  float etadt2[8];
  float tB[8][3], sB[8][3];
  float ve[3];
};

// AdvanceVelocity.f90:39-121 in FP32: CIC gather of E (float4 per node) and
// the Boris rotation in the ve frame (MODE 0); the output-step pair
// AdvanceVelocityImp (MODE 1, :240-347) and AdvanceVelocityExp (MODE 2,
// :128-238) with their norm-preserving rotation.
template <int MODE>
__global__ void __launch_bounds__(kT) k_sp_boris(ParticlesSplit p, int64_t n, GridDims g,
                                                  const float4 *__restrict__ E, BorisF32 bc) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= n) return;
  const int i = p.c[0][q], j = p.c[1][q], k = p.c[2][q];
  if (i < 0 || j < 0 || k < 0 || i > g.nx || j > g.ny || k > g.nz) return;
  const float fx = p.f[0][q], fy = p.f[1][q], fz = p.f[2][q];
  const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
  const int64_t X = g.X, XY = (int64_t)g.X * g.Y;
  const int64_t base = (int64_t)i + X * j + XY * k;
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const float w = wx[a] * wy[b] * wz[c];
        const float4 e = E[base + a + X * b + XY * c];
        ex += e.x * w;
        ey += e.y * w;
        ez += e.z * w;
      }
  const int s = p.sp[q] - 1;
  const float h = bc.etadt2[s];
  const float *t = bc.tB[s], *sb = bc.sB[s];
  if (MODE == 0) {
    const float vm1 = p.v[0][q] - bc.ve[0] + h * ex;
    const float vm2 = p.v[1][q] - bc.ve[1] + h * ey;
    const float vm3 = p.v[2][q] - bc.ve[2] + h * ez;
    const float vd1 = vm1 + vm2 * t[2] - vm3 * t[1];
    const float vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
    const float vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
    const float vp1 = vm1 + vd2 * sb[2] - vd3 * sb[1];
    const float vp2 = vm2 + vd3 * sb[0] - vd1 * sb[2];
    const float vp3 = vm3 + vd1 * sb[1] - vd2 * sb[0];
    p.v[0][q] = vp1 + bc.ve[0] + h * ex;
    p.v[1][q] = vp2 + bc.ve[1] + h * ey;
    p.v[2][q] = vp3 + bc.ve[2] + h * ez;
    return;
  }
  float vm1 = p.v[0][q] - bc.ve[0], vm2 = p.v[1][q] - bc.ve[1], vm3 = p.v[2][q] - bc.ve[2];
  if (MODE == 1) {
    vm1 = vm1 + h * ex;
    vm2 = vm2 + h * ey;
    vm3 = vm3 + h * ez;
  }
  float vd1 = vm1 + vm2 * t[2] - vm3 * t[1];
  float vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
  float vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
  const float sd = vd1 * vd1 + vd2 * vd2 + vd3 * vd3;
  if (sd > 0.0f) {
    const float f = sqrtf((vm1 * vm1 + vm2 * vm2 + vm3 * vm3) / sd);
    vd1 *= f;
    vd2 *= f;
    vd3 *= f;
  }
  if (MODE == 1) {
    p.v[0][q] = vd1 + bc.ve[0];
    p.v[1][q] = vd2 + bc.ve[1];
    p.v[2][q] = vd3 + bc.ve[2];
  } else {
    p.v[0][q] = vd1 + bc.ve[0] + h * ex;
    p.v[1][q] = vd2 + bc.ve[1] + h * ey;
    p.v[2][q] = vd3 + bc.ve[2] + h * ez;
  }
}

// Host FP64 AoS <-> split.
__global__ void __launch_bounds__(kT) k_sp_from_aos(const double *__restrict__ stage,
                                                     const int32_t *__restrict__ stage_sp,
                                                     int64_t count, ParticlesSplit p, int64_t at,
                                                     GridDims g, uint32_t id0) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= count) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xi = (stage[8 * q + d] - g.lmin[d]) / g.dxyz[d] + 0.5;
    double fl = floor(xi);
    float f = (float)(xi - fl);
    if (f >= 1.0f) {
      f = 0.0f;
      fl += 1.0;
    }
    p.c[d][at + q] = (int16_t)fl;
    p.f[d][at + q] = f;
    p.v[d][at + q] = (float)stage[8 * q + 3 + d];
  }
  p.sp[at + q] = (uint8_t)stage_sp[q];
  p.id[at + q] = id0 + (uint32_t)q;
}

__global__ void __launch_bounds__(kT) k_sp_to_aos(ParticlesSplit p, int64_t first, int64_t count,
                                                   GridDims g, double *__restrict__ stage,
                                                   int32_t *__restrict__ stage_sp) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= count) return;
  const int64_t s = first + q;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xi = (double)p.c[d][s] + (double)p.f[d][s];
    stage[8 * q + d] = g.lmin[d] + (xi - 0.5) * g.dxyz[d];
    stage[8 * q + 3 + d] = (double)p.v[d][s];
  }
  stage[8 * q + 6] = 0.0;
  stage[8 * q + 7] = (double)p.id[s];
  stage_sp[q] = p.sp[s];
}

}  // namespace

// ---------------------------------------------------------------------------

int sp_sort(picard_ctx *c, int wrap) {
  const GridDims &g = c->gd;
  const int64_t n = c->n;
  const int64_t K = c->K * g.S;           // (cell, species) keys; bucket K: out-of-grid keepers
  PICARD_HIP(hipMemsetAsync(c->s.hist, 0, sizeof(int32_t) * (K + 2), c->stream));
  PICARD_HIP(hipMemsetAsync(c->s.cursor, 0, sizeof(int32_t) * (K + 1), c->stream));
  if (n > 0) {
    k_sp_flag<<<grid_blocks(n, kT), kT, 0, c->stream>>>(c->ps, n, g, wrap, c->s.key, c->s.hist);
    PICARD_HIP(hipGetLastError());
  }
  PICARD_TRY(scan_exclusive_i32(c, c->s.hist, c->s.hist, K + 1));
  if (n > 0) {
    k_sp_scatter<<<grid_blocks(n, kT), kT, 0, c->stream>>>(c->ps, c->ps2, n, c->s.key, c->s.hist,
                                                           c->s.cursor);
    PICARD_HIP(hipGetLastError());
  }
  PICARD_HIP(hipMemcpyAsync(c->host_counters, c->s.hist + K + 1, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  const ParticlesSplit tmp = c->ps;
  c->ps = c->ps2;
  c->ps2 = tmp;
  c->n = c->host_counters[0];
  return PICARD_OK;
}

int sp_push(picard_ctx *c, double dt) {
  if (c->n <= 0) return PICARD_OK;
  const float3 dtdx = make_float3((float)(dt / c->gd.dxyz[0]), (float)(dt / c->gd.dxyz[1]),
                                  (float)(dt / c->gd.dxyz[2]));
  k_sp_push<<<grid_blocks(c->n, kT), kT, 0, c->stream>>>(c->ps, c->n, dtdx);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int sp_gridify(picard_ctx *c, bool flux) {
  const GridDims &g = c->gd;
  PICARD_HIP(hipMemsetAsync(c->Nfix, 0, sizeof(unsigned long long) * g.S * g.G, c->stream));
  PICARD_HIP(hipMemsetAsync(c->s.counters, 0, sizeof(int32_t), c->stream));
  if (flux) {
    PICARD_TRY(ensure_flux(c));
    PICARD_HIP(hipMemsetAsync(c->Ffix, 0, sizeof(unsigned long long) * 3 * g.S * g.G, c->stream));
  }
  if (c->n > 0) {
    if (flux)
      k_sp_deposit<true><<<grid_blocks((c->n + kDepChunk - 1) / kDepChunk, kT), kT, 0,
                           c->stream>>>(c->ps, c->n, g, c->Nfix, c->Ffix, c->s.counters);
    else
      k_sp_deposit<false><<<grid_blocks((c->n + kDepChunk - 1) / kDepChunk, kT), kT, 0,
                            c->stream>>>(c->ps, c->n, g, c->Nfix, nullptr, c->s.counters);
    PICARD_HIP(hipGetLastError());
  }
  k_sp_fold<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->Nfix, c->N, c->C, c->consts);
  if (flux) k_sp_fold_flux<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->Ffix, c->F, c->consts);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipMemcpyAsync(c->host_counters, c->s.counters, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (c->host_counters[0] != 0) {
    set_error("ParticlesGridify (split): " + std::to_string(c->host_counters[0]) +
              " particle(s) outside the grid were not deposited");
    return PICARD_ERR_RANGE;
  }
  return PICARD_OK;
}

// One sweep launch (timed with events t0/t1 when given).
template <bool PX>
void launch_sweep(bool pad, int blocks, hipStream_t s, hipEvent_t t0, hipEvent_t t1,
                  unsigned long long *ks,
                  const SweepGeo &sg, const SweepCoef &cf, const float *r32, const float *e,
                  float *out, float *eo_out, const float *eo_prev, const double *pin, double *pout,
                  JacobiState *js, const MxLaunch &L) {
  auto fn = pad ? k_mx_sweep<PX, true> : k_mx_sweep<PX, false>;
  hipExtLaunchKernelGGL(fn, dim3(blocks), dim3(64, kSwRows), 0, s, t0, t1, 0, sg, cf, r32, e,
                        out, eo_out, eo_prev, pin, pout, js, L, ks);
}

// Resident blocks per XCD of the sweep kernel (the heaviest of the solve).
int sweep_wmax(picard_ctx *c) {
  if (c->sweep_wmax <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mx_sweep<false, false>, kSwThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8) {
      c->sweep_wmax = 256;
    } else {
      c->sweep_wmax = (cus / 8) * per_cu;
    }
  }
  return c->sweep_wmax;
}

// The persistent schedule of the outer-step kernels, with no more blocks per
// XCD than are resident at once, so that the XCD slab assignment is not split
// over two dispatch rounds.
TileSched stencil_sched(picard_ctx *c) {
  const GridDims &g = c->gd;
  TileSched ts = make_sched(g.nx, g.ny, g.nz, kStencilT);
  const int wmax = sweep_wmax(c);
  if (ts.W > wmax) {
    ts.W = wmax < 4 ? 4 : wmax;
    ts.B = 8 * ts.W;
  }
  return ts;
}

// Tiles of the sweep kernel: row segments x 4-row groups x z chunks of an
// XCD slab.  The slab is cut into chunks only until the tiles give about half
// the resident blocks per XCD (at 256^3 and 4 blocks per CU: one chunk, 512
// blocks of 4 waves marching 32 planes, measured 31 us against 38 us for two
// chunks and 1024 blocks).
// rows: rows per wave; fill: z chunks until the tiles fill the resident
// blocks (else about half of them); seg: cells of a row a wave keeps
SweepGeo sweep_geo(picard_ctx *c, int wmax, int rows = 1, bool fill = false, int seg = kSwSeg) {
  const GridDims &g = c->gd;
  SweepGeo sg;
  sg.nx = g.nx;
  sg.ny = g.ny;
  sg.nz = g.nz;
  sg.NX = (uint32_t)(g.nx + 3) & ~3u;
  sg.P = sg.NX * (uint32_t)g.ny;
  sg.px = !g.flow[0];
  sg.py = !g.flow[1];
  sg.pz = !g.flow[2];
  sg.ntx = (g.nx + seg - 1) / seg;
  sg.nty = (g.ny + kSwRows * rows - 1) / (kSwRows * rows);
  const int per_chunk = sg.ntx * sg.nty;
  const int span = fill ? per_chunk : 2 * per_chunk;
  int nzc = (wmax + span - 1) / span;
  const int zmax = g.nz / 64 > 1 ? g.nz / 64 : 1;         // chunks of at least 8 planes
  if (nzc > zmax) nzc = zmax;
  if (nzc < 1) nzc = 1;
  sg.nzc = nzc;
  sg.TP = per_chunk * nzc;
  // equal tiles per block: the fewest blocks that need no more passes
  int W = sg.TP < wmax ? sg.TP : wmax;
  W = (sg.TP + (sg.TP + W - 1) / W - 1) / ((sg.TP + W - 1) / W);
  if (W > (int)(c->s.partials_len / 8)) W = (int)(c->s.partials_len / 8);
  sg.W = W < 1 ? 1 : W;
  return sg;
}

SweepGeo sweep_sched(picard_ctx *c) { return sweep_geo(c, sweep_wmax(c)); }

// The two-sweep kernel's schedule: the same tiles, sized by its own
// occupancy (it holds more registers per wave); rows wider than one segment
// in overlapping segments that keep kPairKeep cells each.
SweepGeo pair_sched(picard_ctx *c) {
  if (c->pair_wmax <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mx_pair<false, false, kPairRows, false, false>, kSwThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8) {
      c->pair_wmax = 64;
    } else {
      c->pair_wmax = (cus / 8) * per_cu;
    }
  }
  return sweep_geo(c, c->pair_wmax, kPairRows, true, c->gd.nx > kSwSeg ? kPairKeep : kSwSeg);
}

// The LDS two-sweep kernel's schedule: tiles of at most kLdsRows rows (the
// rows split evenly over ceil(ny / kLdsRows) tiles) x z chunks of an XCD
// slab, as many chunks as fill the resident blocks (chunks of at least 8
// planes), blocks per XCD no more than are resident.
SweepGeo lds_sched(picard_ctx *c) {
  const GridDims &g = c->gd;
  if (c->lds_wmax <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_mx_pair_lds<false, false, kLdsWaves, kLdsDepth>, 64 * kLdsWaves, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8) {
      c->lds_wmax = 32;
    } else {
      c->lds_wmax = (cus / 8) * per_cu;
    }
  }
  const int wmax = c->lds_wmax;
  SweepGeo sg;
  sg.nx = g.nx;
  sg.ny = g.ny;
  sg.nz = g.nz;
  sg.NX = (uint32_t)(g.nx + 3) & ~3u;
  sg.P = sg.NX * (uint32_t)g.ny;
  sg.px = !g.flow[0];
  sg.py = !g.flow[1];
  sg.pz = !g.flow[2];
  sg.ntx = 1;
  sg.nty = (g.ny + kLdsRows - 1) / kLdsRows;
  int nzc = wmax / sg.nty;
  const int zmax = g.nz / 64 > 1 ? g.nz / 64 : 1;
  if (nzc > zmax) nzc = zmax;
  if (nzc < 1) nzc = 1;
  sg.nzc = nzc;
  sg.TP = sg.nty * nzc;
  int W = sg.TP < wmax ? sg.TP : wmax;
  W = (sg.TP + (sg.TP + W - 1) / W - 1) / ((sg.TP + W - 1) / W);
  if (W > (int)(c->s.partials_len / 8)) W = (int)(c->s.partials_len / 8);
  sg.W = W < 1 ? 1 : W;
  return sg;
}

void launch_pair_lds(bool pad, int blocks, hipStream_t s, hipEvent_t t0, hipEvent_t t1,
                  unsigned long long *ks,
                     const SweepGeo &sg, const SweepCoef &cf, const float *r32, const float *e,
                     float *out, float *eo_out, const float *eo_prev, const double *pin,
                     double *pout, JacobiState *js, const MxLaunch &L) {
  constexpr int W = kLdsWaves, Dp = kLdsDepth;
  auto fn = L.j == 0 ? (pad ? k_mx_pair_lds<true, true, W, Dp> : k_mx_pair_lds<false, true, W, Dp>)
                     : (pad ? k_mx_pair_lds<true, false, W, Dp> : k_mx_pair_lds<false, false, W, Dp>);
  hipExtLaunchKernelGGL(fn, dim3(blocks), dim3(64, kLdsWaves), 0, s, t0, t1, 0, sg, cf, r32, e,
                        out, eo_out, eo_prev, pin, pout, js, L, ks);
}

// The three-sweep LDS kernel's schedule: tiles of at most kTriRows rows x z
// chunks of an XCD slab, as lds_sched.
SweepGeo tri_sched(picard_ctx *c) {
  const GridDims &g = c->gd;
  if (c->tri_wmax <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_mx_triple_lds<false, false, kTriWaves>, 64 * kTriWaves, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8) {
      c->tri_wmax = 32;
    } else {
      c->tri_wmax = (cus / 8) * per_cu;
    }
  }
  const int wmax = c->tri_wmax;
  SweepGeo sg;
  sg.nx = g.nx;
  sg.ny = g.ny;
  sg.nz = g.nz;
  sg.NX = (uint32_t)(g.nx + 3) & ~3u;
  sg.P = sg.NX * (uint32_t)g.ny;
  sg.px = !g.flow[0];
  sg.py = !g.flow[1];
  sg.pz = !g.flow[2];
  sg.ntx = 1;
  sg.nty = (g.ny + kTriRows - 1) / kTriRows;
  int nzc = wmax / sg.nty;
  const int zmax = g.nz / 64 > 1 ? g.nz / 64 : 1;
  if (nzc > zmax) nzc = zmax;
  if (nzc < 1) nzc = 1;
  sg.nzc = nzc;
  sg.TP = sg.nty * nzc;
  int W = sg.TP < wmax ? sg.TP : wmax;
  W = (sg.TP + (sg.TP + W - 1) / W - 1) / ((sg.TP + W - 1) / W);
  if (W > (int)(c->s.partials_len / 8)) W = (int)(c->s.partials_len / 8);
  sg.W = W < 1 ? 1 : W;
  return sg;
}

void launch_triple_lds(bool pad, int blocks, hipStream_t s, hipEvent_t t0, hipEvent_t t1,
                  unsigned long long *ks,
                       const SweepGeo &sg, const SweepCoef &cf, const float *r32, const float *e,
                       float *out, float *eo_out, const float *eo_prev, const double *pin,
                       double *pout, JacobiState *js, const MxLaunch &L) {
  constexpr int W = kTriWaves;
  auto fn = L.j == 0 ? (pad ? k_mx_triple_lds<true, true, W> : k_mx_triple_lds<false, true, W>)
                     : (pad ? k_mx_triple_lds<true, false, W> : k_mx_triple_lds<false, false, W>);
  hipExtLaunchKernelGGL(fn, dim3(blocks), dim3(64, kTriWaves), 0, s, t0, t1, 0, sg, cf, r32, e,
                        out, eo_out, eo_prev, pin, pout, js, L, ks);
}

// One two-sweep launch (timed with events t0/t1 when given).
template <bool PX, bool MULTI>
void launch_pair(bool pad, int blocks, hipStream_t s, hipEvent_t t0, hipEvent_t t1,
                  unsigned long long *ks,
                 const SweepGeo &sg, const SweepCoef &cf, const float *r32, const float *e,
                 float *out, float *eo_out, const float *eo_prev, const double *pin, double *pout,
                 JacobiState *js, const MxLaunch &L) {
  constexpr bool M = MULTI;
  auto fn = L.j == 0 ? (pad ? k_mx_pair<PX, true, kPairRows, true, M> : k_mx_pair<PX, false, kPairRows, true, M>)
                     : (pad ? k_mx_pair<PX, true, kPairRows, false, M> : k_mx_pair<PX, false, kPairRows, false, M>);
  hipExtLaunchKernelGGL(fn, dim3(blocks), dim3(64, kSwRows), 0, s, t0, t1, 0, sg, cf, r32, e,
                        out, eo_out, eo_prev, pin, pout, js, L, ks);
}

// Tiles of the outer-step kernel: 64-cell row segments x 4-row groups x z
// chunks of an XCD slab, about half the resident blocks per XCD, as the sweep.
OuterGeo outer_sched(picard_ctx *c) {
  const GridDims &g = c->gd;
  if (c->outer_wmax <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mx_outer<true>, kStencilT, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8) {
      c->outer_wmax = 128;
    } else {
      c->outer_wmax = (cus / 8) * per_cu;
    }
  }
  OuterGeo og;
  og.nx = g.nx;
  og.ny = g.ny;
  og.nz = g.nz;
  og.X = (uint32_t)g.X;
  og.XY = (uint32_t)g.X * (uint32_t)g.Y;
  og.NX = (uint32_t)(g.nx + 3) & ~3u;
  og.P = og.NX * (uint32_t)g.ny;
  og.px = !g.flow[0];
  og.py = !g.flow[1];
  og.pz = !g.flow[2];
  og.ntx = (g.nx + 63) / 64;
  og.nty = (g.ny + kOutRows - 1) / kOutRows;
  const int wmax = c->outer_wmax;
  const int per_chunk = og.ntx * og.nty;
  int nzc = (wmax + 2 * per_chunk - 1) / (2 * per_chunk);
  const int zmax = g.nz / 64 > 1 ? g.nz / 64 : 1;
  if (nzc > zmax) nzc = zmax;
  if (nzc < 1) nzc = 1;
  og.nzc = nzc;
  og.TP = per_chunk * nzc;
  // equal tiles per block: the fewest blocks that need no more passes
  int W = og.TP < wmax ? og.TP : wmax;
  W = (og.TP + (og.TP + W - 1) / W - 1) / ((og.TP + W - 1) / W);
  if (W > (int)(c->s.partials_len / 8)) W = (int)(c->s.partials_len / 8);
  og.W = W < 1 ? 1 : W;
  return og;
}

int sp_epotential(picard_ctx *c, double maxerr, int32_t maxit, int64_t *sweeps) {
  const GridDims &g = c->gd;
  const double dx = g.dxyz[0], dy = g.dxyz[1], dz = g.dxyz[2];
  const double den = 1.0 / ((dx * dy) * (dx * dy) + (dx * dz) * (dx * dz) + (dy * dz) * (dy * dz));
  MixCoef mc;
  mc.ax = 0.5 * den * ((dy * dz) * (dy * dz));
  mc.ay = 0.5 * den * ((dx * dz) * (dx * dz));
  mc.az = 0.5 * den * ((dx * dy) * (dx * dy));
  mc.af = 0.5 * den * ((dx * dy * dz) * (dx * dy * dz));
  mc.dx2 = dx * dx;
  mc.dy2 = dy * dy;
  mc.dz2 = dz * dz;
  mc.axf = (float)mc.ax; mc.ayf = (float)mc.ay; mc.azf = (float)mc.az; mc.aff = (float)mc.af;
  mc.idx2f = (float)(1.0 / mc.dx2); mc.idy2f = (float)(1.0 / mc.dy2); mc.idz2f = (float)(1.0 / mc.dz2);
  JacobiState init{};
  init.maxit = maxit;
  init.last = 2 * (int64_t)maxit + 1;
  PICARD_HIP(hipMemcpyAsync(c->js, &init, sizeof(JacobiState), hipMemcpyHostToDevice, c->stream));
  PICARD_TRY(launch_bnorm(c, maxerr));      // launch 0: b, |b|, the test of ss = 1
  const TileSched ts = stencil_sched(c);
  const SweepGeo sg = sweep_sched(c);
  const OuterGeo og = outer_sched(c);
  const CompactDims cd{sg.NX, sg.P};
  const int swB = 8 * sg.W, ouB = 8 * og.W;
  const bool pad = (g.nx & 3) != 0;
  // u ping-pongs between U and the work buffer (outer step s reads Ub[s % 2]
  // and writes Ub[(s + 1) % 2]); both carry the boundary values in the guards
  double *Ub[2] = {c->U, c->Uw};
  PICARD_HIP(hipMemcpyAsync(c->Uw, c->U, sizeof(double) * g.G, hipMemcpyDeviceToDevice, c->stream));
  int outer = 0;
  SweepCoef cf{};
  {
    // FP32 weights with 2 (wx + wy + wz) = 1 as closely as FP32 allows
    cf.wx = (float)mc.ax;
    cf.wy = (float)mc.ay;
    float best = (float)(0.5 - (double)cf.wx - (double)cf.wy);
    for (float cand : {best, nextafterf(best, 1.0f), nextafterf(best, 0.0f)}) {
      const double err = std::fabs(0.5 - ((double)cf.wx + (double)cf.wy + (double)cand));
      const double cur = std::fabs(0.5 - ((double)cf.wx + (double)cf.wy + (double)best));
      if (err < cur) best = cand;
    }
    cf.wz = best;
    cf.kappa = (float)(1.0 - 2.0 * ((double)cf.wx + (double)cf.wy + (double)cf.wz));
    cf.af = mc.aff;
    cf.idx2 = mc.idx2f;
    cf.idy2 = mc.idy2f;
    cf.idz2 = mc.idz2f;
  }
  const int m = c->mx_inner;
  // two sweeps per launch: rows of one segment, or wider rows with x not
  // periodic (overlapping segments), and an even inner length
  const bool pairs = c->mx_pairs && (sg.ntx == 1 || g.flow[0]) && m % 2 == 0;
  // rows of one segment with x not periodic: rows shared through LDS, three
  // sweeps per launch (or two where the sweeps left call for it), except in
  // the launches that test the stop, which take the register-window
  // two-sweep kernel (it evaluates the residual in its first sweep, 3 us
  // more per launch, where the LDS kernels need a pass of their own, 30 us):
  // the launches the host marks, and every launch of a refinement step that
  // begins with the stop near (the flag the last one left; near the end of a
  // slowly converging solve it stays set for many refinement steps)
  const bool lds = pairs && c->pair_lds && sg.ntx == 1 && g.flow[0];
  const bool tri = lds && c->mx_triples;
  const SweepGeo pgeo = pairs ? pair_sched(c) : sg;
  const SweepGeo lgeo = lds ? lds_sched(c) : pgeo;
  const SweepGeo tgeo = tri ? tri_sched(c) : pgeo;
  const int prB = 8 * pgeo.W, plB = 8 * lgeo.W, ptB = 8 * tgeo.W;
  double *P[2] = {c->s.partials, c->s.partials + c->s.partials_len};
  float *E0 = c->e32[0], *E1 = c->e32[1];
  float *const E[2] = {E0, E1};
  // a launch reads e_j from E[cur] and writes its last iterate to the other
  // buffer; the launch input the stop test picks, and e_m, are found by the
  // buffer index the host passes (ibuf, cbuf) and the stop slot keeps
  int cur = 0;
  const int64_t last_it = init.last;
  // boundary copies of e_k in slot k % 4: a launch of s sweeps with input j
  // reads those of e_{j-1} (its stop test) and writes those of e_{j+s-1}
  // (e_{j+1} for a single sweep), which the next launch's test takes
  float *const *EO = c->eo32;
  int32_t seq = 1, base = 0;
  int pw = 0;                               // partials buffer the next writer uses
  int32_t check = -1;                       // iterate whose residual the last writer stored
  int32_t cbuf = 0;                         // the buffer holding iterate `check`
  bool prevtest = true;                     // the last launch tested whatever its near flag
  int32_t nin = 0;                          // partials the last writer stored
  StopSlot fin{};
  bool stop_near = false;                   // the stop test found the stop near
  while (true) {
    // timing: one launch in kTimeEvery, with dispatch-packet events around it
    // and the kernel's own clock stamps (events around every launch would
    // slow the timed solve itself)
    const int timed = (m + kTimeEvery - 1) / kTimeEvery;         // one launch position in 16
    if (c->timing) {
      while ((int64_t)c->tev.size() < 2 * timed) {
        hipEvent_t e;
        PICARD_HIP(hipEventCreate(&e));
        c->tev.push_back(e);
      }
      PICARD_TRY(kstamp_reset(c, timed));
    }
    // outer residual of iterate `base` (its stop test happens in the next
    // launch); the stale-halo norm needs e_m and e_{m-1} of the previous outer
    double *usrc = Ub[outer & 1], *udst = Ub[(outer + 1) & 1];
    {
      MxLaunch L{seq, base, 0, -1, m, 0, pairs ? 1 : 0, 1, 1, cur, cur};
      // e_m and the boundary copies of e_{m-1} of the previous outer step
      const float *em = E[cur];
      const float *eo = EO[(m - 1) % 4];
      if (base > 0)
        k_mx_outer<true><<<ouB, dim3(64, kOutRows), 0, c->stream>>>(
            og, mc.dx2, mc.dy2, mc.dz2, usrc, udst, c->b, c->r32, em, eo, P[pw ^ 1], P[pw],
            c->js, L);
      else
        k_mx_outer<false><<<ouB, dim3(64, kOutRows), 0, c->stream>>>(
            og, mc.dx2, mc.dy2, mc.dz2, usrc, udst, c->b, c->r32, nullptr, nullptr, P[pw ^ 1],
            P[pw], c->js, L);
      ++seq;
      check = base;
      cbuf = cur;
      nin = ouB;
      pw ^= 1;
      prevtest = true;
    }
    int nrec = 0, tsw = 0;                  // timed launches of this outer step, their sweeps
    int li = 0;                             // launches of this outer step
    int jt = 2 * kPairTestEvery;            // the next iterate position due a stop test
    for (int j = 0; j < m; ++li) {
      // the iterate-0 test used b itself (launch 0); later ones use residuals
      const int32_t chk = (check >= 1) ? check : -1;
      const int r = m - j;                  // sweeps left in this refinement step
      const int64_t d = last_it - ((int64_t)base + j);   // sweeps to iterate 2*maxit+1
      // a launch of s sweeps may not pass over iterate 2*maxit+1 (the
      // solve's last, which must exist) and leaves no single sweep before the
      // end of the refinement step (the next one's residual needs e_{m-1})
      auto fits = [&](int sz) { return sz <= r && (d >= sz || d <= 0); };
      // the stop test of iterate base + j: every single sweep, about every
      // 2 kPairTestEvery iterates with more per launch and, on the device,
      // every launch while the stop is near
      const bool due = j >= 1 && j >= jt;
      int size = 1;
      if (pairs) {
        if (tri && fits(3) && r != 4 && (!due || r == 3) && !stop_near) size = 3;
        else if (fits(2)) size = 2;
      }
      const bool test = j >= 1 && (size == 1 || due);
      if (test && size > 1) jt = j + 2 * kPairTestEvery;
      MxLaunch L{seq, base, j, chk, m, nin, pairs ? 1 : 0, test ? 1 : 0, prevtest ? 1 : 0,
                 cur, cbuf};
      const double *pin = P[pw ^ 1];
      double *pout = P[pw];
      // e_j (none for a single first sweep) -> e_{j+size}
      const float *e = j >= 1 || size > 1 ? E[cur] : nullptr;
      float *out = E[cur ^ 1];
      float *eo_out = EO[(size == 1 ? j + 1 : j + size - 1) % 4];
      const float *eo_prev = j >= 2 ? EO[(j - 1) % 4] : nullptr;
      // the timed position moves by two launches per refinement step, so that
      // launches with and without the stop test are sampled in proportion
      const int tpos = (kTimeEvery / 2 + 2 * outer) % kTimeEvery;
      const bool tj = c->timing && j >= 1 && li % kTimeEvery == tpos && nrec < timed;
      unsigned long long *ks = tj ? c->kst + picard_ctx::kStampWords * nrec : nullptr;
      hipEvent_t t0 = tj ? c->tev[2 * nrec] : nullptr;
      hipEvent_t t1 = tj ? c->tev[2 * nrec + 1] : nullptr;
      if (tj) {
        ++nrec;
        tsw += size;
      }
      const bool use_lds = size == 2 && lds && !test && !stop_near;
      int blocks = swB;
      if (size == 3) {
        blocks = ptB;
        launch_triple_lds(pad, ptB, c->stream, t0, t1, ks, tgeo, cf, c->r32, e, out, eo_out, eo_prev,
                          pin, pout, c->js, L);
      } else if (use_lds) {
        blocks = plB;
        launch_pair_lds(pad, plB, c->stream, t0, t1, ks, lgeo, cf, c->r32, e, out, eo_out, eo_prev,
                        pin, pout, c->js, L);
      } else if (size == 2) {
        blocks = prB;
        if (sg.px)
          launch_pair<true, false>(pad, prB, c->stream, t0, t1, ks, pgeo, cf, c->r32, e, out, eo_out,
                                   eo_prev, pin, pout, c->js, L);
        else if (pgeo.ntx > 1)
          launch_pair<false, true>(pad, prB, c->stream, t0, t1, ks, pgeo, cf, c->r32, e, out, eo_out,
                                   eo_prev, pin, pout, c->js, L);
        else
          launch_pair<false, false>(pad, prB, c->stream, t0, t1, ks, pgeo, cf, c->r32, e, out,
                                    eo_out, eo_prev, pin, pout, c->js, L);
      } else if (sg.px) {
        launch_sweep<true>(pad, swB, c->stream, t0, t1, ks, sg, cf, c->r32, e, out, eo_out, eo_prev,
                           pin, pout, c->js, L);
      } else {
        launch_sweep<false>(pad, swB, c->stream, t0, t1, ks, sg, cf, c->r32, e, out, eo_out, eo_prev,
                            pin, pout, c->js, L);
      }
      ++seq;
      if (j >= 1) {               // may have written the residual of iterate base + j
        check = base + j;
        cbuf = cur;
        nin = blocks;
        pw ^= 1;
      } else {
        check = -1;
      }
      prevtest = test;
      cur ^= 1;
      j += size;
    }
    {
      const int32_t chk = (check >= 1) ? check : -1;
      MxLaunch L{seq, base, 0, chk, m, nin, pairs ? 1 : 0, 1, prevtest ? 1 : 0, cur, cbuf};
      k_mx_update<<<ts.B, dim3(ts.bx, ts.by), 0, c->stream>>>(g, cd, ts, udst, E0, E1, P[pw ^ 1],
                                                              c->js, L);
      ++seq;
      check = -1;
    }
    PICARD_HIP(hipGetLastError());
    PICARD_HIP(hipMemcpyAsync(c->js_host, c->js, sizeof(JacobiState), hipMemcpyDeviceToHost,
                              c->stream));
    PICARD_HIP(hipStreamSynchronize(c->stream));
    if (c->timing) {
      for (int q = 0; q < nrec; ++q) {
        float ms = 0.0f;
        PICARD_HIP(hipEventElapsedTime(&ms, c->tev[2 * q], c->tev[2 * q + 1]));
        c->t_ms += ms;
      }
      c->t_launches += nrec;
      c->t_sweeps += tsw;
      PICARD_TRY(kstamp_collect(c, nrec));
    }
    fin = c->js_host->slot[(seq - 1) & 1];
    stop_near = fin.near != 0;
    if (fin.done) {
      if (udst != c->U)
        PICARD_HIP(hipMemcpyAsync(c->U, udst, sizeof(double) * g.G, hipMemcpyDeviceToDevice,
                                  c->stream));
      break;
    }
    base += m;
    ++outer;
  }
  k_mx_halo<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->U);
  PICARD_HIP(hipGetLastError());
  *sweeps = fin.result;
  c->last_sweeps = fin.result;
  return fin.result == init.last ? PICARD_WARN_NOCONV : PICARD_OK;
}

int sp_efield(picard_ctx *c) {
  k_sp_efield<<<grid_blocks(c->gd.G, kT), kT, 0, c->stream>>>(c->gd, c->U, c->E32);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int sp_boris(picard_ctx *c, double dt, int mode, double deltat) {
  if (c->n <= 0) return PICARD_OK;
  BorisF32 bc;
  const double scale = mode == 0 ? 1.0 : 2.0 * deltat / dt;   // AdvanceVelocity.f90:156-157
  for (int s = 0; s < 8; ++s) {
    bc.etadt2[s] = (float)(0.5 * dt * c->spc.eta[s] * scale);
    for (int q = 0; q < 3; ++q) {
      bc.tB[s][q] = (float)(c->spc.tB[4 * s + q] * scale);
      bc.sB[s][q] = (float)c->spc.sB[4 * s + q];
    }
  }
  for (int q = 0; q < 3; ++q) bc.ve[q] = (float)c->spc.ve[q];
  const int nb = grid_blocks(c->n, kT);
  const float4 *E4 = reinterpret_cast<const float4 *>(c->E32);
  if (mode == 0) k_sp_boris<0><<<nb, kT, 0, c->stream>>>(c->ps, c->n, c->gd, E4, bc);
  else if (mode == 1) k_sp_boris<1><<<nb, kT, 0, c->stream>>>(c->ps, c->n, c->gd, E4, bc);
  else k_sp_boris<2><<<nb, kT, 0, c->stream>>>(c->ps, c->n, c->gd, E4, bc);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int sp_append(picard_ctx *c, const double *host_rp, const int32_t *host_sp, int64_t count) {
  if (count <= 0) return PICARD_OK;
  if (c->n + count > c->cap) {
    set_error("particles_append: capacity max_particles exceeded");
    return PICARD_ERR_CAPACITY;
  }
  const int64_t chunk = int64_t(1) << 21;
  PICARD_TRY(ensure_stage(c, count < chunk ? count : chunk));
  for (int64_t done = 0; done < count; done += chunk) {
    const int64_t m = (count - done) < chunk ? (count - done) : chunk;
    PICARD_HIP(hipMemcpyAsync(c->s.stage, host_rp + 8 * done, sizeof(double) * 8 * m,
                              hipMemcpyHostToDevice, c->stream));
    PICARD_HIP(hipMemcpyAsync(c->s.stage_sp, host_sp + done, sizeof(int32_t) * m,
                              hipMemcpyHostToDevice, c->stream));
    k_sp_from_aos<<<grid_blocks(m, kT), kT, 0, c->stream>>>(c->s.stage, c->s.stage_sp, m, c->ps,
                                                            c->n + done, c->gd,
                                                            (uint32_t)(c->next_id + done));
    PICARD_HIP(hipGetLastError());
    PICARD_HIP(hipStreamSynchronize(c->stream));
  }
  c->n += count;
  c->next_id += count;
  return PICARD_OK;
}

int sp_stage(picard_ctx *c, int64_t first, int64_t count, double *stage, int32_t *stage_sp) {
  if (count <= 0) return PICARD_OK;
  k_sp_to_aos<<<grid_blocks(count, kT), kT, 0, c->stream>>>(c->ps, first, count, c->gd, stage,
                                                            stage_sp);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int sp_read(picard_ctx *c, int64_t first, int64_t count, double *host_rp, int32_t *host_sp) {
  if (count <= 0) return PICARD_OK;
  if (first < 0 || first + count > c->n) {
    set_error("particles_read: range outside [0, num_local)");
    return PICARD_ERR_ARG;
  }
  const int64_t chunk = int64_t(1) << 21;
  PICARD_TRY(ensure_stage(c, count < chunk ? count : chunk));
  for (int64_t done = 0; done < count; done += chunk) {
    const int64_t m = (count - done) < chunk ? (count - done) : chunk;
    k_sp_to_aos<<<grid_blocks(m, kT), kT, 0, c->stream>>>(c->ps, first + done, m, c->gd, c->s.stage,
                                                          c->s.stage_sp);
    PICARD_HIP(hipGetLastError());
    PICARD_HIP(hipMemcpyAsync(host_rp + 8 * done, c->s.stage, sizeof(double) * 8 * m,
                              hipMemcpyDeviceToHost, c->stream));
    if (host_sp)
      PICARD_HIP(hipMemcpyAsync(host_sp + done, c->s.stage_sp, sizeof(int32_t) * m,
                                hipMemcpyDeviceToHost, c->stream));
    PICARD_HIP(hipStreamSynchronize(c->stream));
  }
  return PICARD_OK;
}

// E as FP64 (3 per node) for field_read in the split form.
namespace {
__global__ void k_sp_e_widen(int64_t G, const float *__restrict__ E32, double *__restrict__ E) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= G) return;
  for (int d = 0; d < 3; ++d) E[3 * node + d] = (double)E32[4 * node + d];
}
}  // namespace

int sp_widen_e(picard_ctx *c) {
  k_sp_e_widen<<<grid_blocks(c->gd.G, kT), kT, 0, c->stream>>>(c->gd.G, c->E32, c->E);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

}  // namespace picard
