// grid.hip -- FP64 grid kernels: ParticlesGridify (reference-order CIC
// deposit), CalcDistance, BCpotentials, CalcEpotential (Jacobi with a
// device-side stop), CalcEfield and BCfields, for one domain.
//
// One domain means every MPI exchange of the reference is a self-copy; those
// copies become index arithmetic here (guard folds, periodic wraps), with the
// same z -> y -> x composition order.
#include <hip/hip_ext.h>

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

// ---------------------------------------------------------------------------
// ParticlesGridify (ParticlesGridify.f90:56-293)
//
// The reference accumulates N(s, node) sequentially over pp = n..1.  To get the
// same sums bit for bit in parallel, particles are bucketed by (species, lower
// node) and every (species, node) output merges its 8 contributing buckets in
// descending particle index, adding in exactly the reference's order.

__device__ __forceinline__ int64_t cell_key(const GridDims &g, int s0, const CicLoc &L) {
  const int64_t K = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  return (int64_t)s0 * K + (L.idx[0] - 1) +
         (int64_t)(g.nx + 1) * ((L.idx[1] - 1) + (int64_t)(g.ny + 1) * (L.idx[2] - 1));
}

__global__ void __launch_bounds__(kT) k_dep_key(ParticlesF64 p, int64_t n, GridDims g,
                                                 int32_t *__restrict__ key,
                                                 int32_t *__restrict__ hist,
                                                 int32_t *__restrict__ errors) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const CicLoc L = cic_locate(g, p.r[0][i], p.r[1][i], p.r[2][i]);
  const int s0 = p.sp[i] - 1;
  if (!cic_in_bounds(g, L) || s0 < 0 || s0 >= g.S) {
    key[i] = -1;
    atomicAdd(errors, 1);
    return;
  }
  const int32_t k = (int32_t)cell_key(g, s0, L);
  key[i] = k;
  atomicAdd(&hist[k], 1);
}

__global__ void __launch_bounds__(kT) k_dep_scatter(int64_t n, const int32_t *__restrict__ key,
                                                     const int32_t *__restrict__ off,
                                                     int32_t *__restrict__ cursor,
                                                     int32_t *__restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int32_t k = key[i];
  if (k < 0) return;
  const int32_t pos = off[k] + atomicAdd(&cursor[k], 1);
  perm[pos] = (int32_t)i;
}

// Each bucket sorted by descending particle index (insertion sort; buckets
// hold a few dozen particles).
__global__ void __launch_bounds__(kT) k_dep_bucket_sort(int64_t nbuckets,
                                                         const int32_t *__restrict__ off,
                                                         int32_t *__restrict__ perm) {
  const int64_t b = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (b >= nbuckets) return;
  const int32_t lo = off[b], hi = off[b + 1];
  for (int32_t a = lo + 1; a < hi; ++a) {
    const int32_t v = perm[a];
    int32_t q = a - 1;
    while (q >= lo && perm[q] < v) {
      perm[q + 1] = perm[q];
      --q;
    }
    perm[q + 1] = v;
  }
}

// One thread per (node, species): 8-way merge by descending particle index.
// FLUX also accumulates the number fluxes of ParticlesGridifyAll
// (ParticlesGridifyAll.f90:126-212), F(s, c, node) += F_k n2p(s) v_c, in the
// same order; Fraw is laid out (s fastest, then c, then node).
template <bool FLUX>
__global__ void __launch_bounds__(kT) k_dep_gather(ParticlesF64 p, GridDims g, const double *n2p,
                                                    const int32_t *__restrict__ off,
                                                    const int32_t *__restrict__ perm,
                                                    double *__restrict__ Nraw,
                                                    double *__restrict__ Fraw) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= g.G * g.S) return;
  const int s0 = (int)(t % g.S);
  const int64_t node = t / g.S;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  const int64_t K = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  int32_t head[8], end[8];
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    const int ci = i - ox, cj = j - oy, ck = k - oz;
    if (ci >= 1 && cj >= 1 && ck >= 1 && ci <= g.nx + 1 && cj <= g.ny + 1 && ck <= g.nz + 1) {
      const int64_t b = (int64_t)s0 * K + (ci - 1) +
                        (int64_t)(g.nx + 1) * ((cj - 1) + (int64_t)(g.ny + 1) * (ck - 1));
      head[o] = off[b];
      end[o] = off[b + 1];
    } else {
      head[o] = 0;
      end[o] = 0;
    }
  }
  const double w = n2p[s0];
  double acc = 0.0;                      // N = 0.0d0 (ParticlesGridify.f90:59)
  double fl[3] = {0.0, 0.0, 0.0};        // F = 0.0d0 (ParticlesGridifyAll.f90:61)
  for (;;) {
    int best = -1;
    int32_t bestp = -1;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      if (head[o] < end[o]) {
        const int32_t pp = perm[head[o]];
        if (pp > bestp) {
          bestp = pp;
          best = o;
        }
      }
    }
    if (best < 0) break;
    ++head[best];
    const CicLoc L = cic_locate(g, p.r[0][bestp], p.r[1][bestp], p.r[2][bestp]);
    // node = idx + (ox, oy, oz): factor is rest where the offset is 1, drest where 0,
    // which reproduces F1..F8 of ParticlesGridify.f90:81-88
    const double fx = (best & 1) ? L.rest[0] : L.drest[0];
    const double fy = (best & 2) ? L.rest[1] : L.drest[1];
    const double fz = (best & 4) ? L.rest[2] : L.drest[2];
    const double F = fx * fy * fz / g.dV;
    acc = acc + F * w;
    if (FLUX) {
      const double Fw = F * w;
      fl[0] = fl[0] + Fw * p.r[3][bestp];
      fl[1] = fl[1] + Fw * p.r[4][bestp];
      fl[2] = fl[2] + Fw * p.r[5][bestp];
    }
  }
  Nraw[node * g.S + s0] = acc;
  if (FLUX)
    for (int c = 0; c < 3; ++c) Fraw[(node * 3 + c) * g.S + s0] = fl[c];
}

// Guard handling of one dimension on one rank (ParticlesGridify.f90:129-285):
// up to two source planes per node; *zero = the reference assigns 0.0d0.
__device__ __forceinline__ int fold_src(int m, int n, int flow, int src[2], bool *zero) {
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

// Folds z, then y, then x (the reference's order), then C = sum_s N(s)*charge(s).
__global__ void __launch_bounds__(kT) k_fold_charge(GridDims g, const double *__restrict__ Nraw,
                                                     double *__restrict__ N,
                                                     double *__restrict__ C,
                                                     const double *charge) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int sx[2], sy[2], sz[2];
  bool zx, zy, zz;
  const int nx_ = fold_src(i, g.nx, g.flow[0], sx, &zx);
  const int ny_ = fold_src(j, g.ny, g.flow[1], sy, &zy);
  const int nz_ = fold_src(k, g.nz, g.flow[2], sz, &zz);
  double cacc = 0.0;                       // C = 0.0d0 (ParticlesGridify.f90:60)
  for (int s = 0; s < g.S; ++s) {
    double vx = 0.0;
    if (!zx) {
      for (int a = 0; a < nx_; ++a) {
        double vy = 0.0;
        if (!zy) {
          for (int bb = 0; bb < ny_; ++bb) {
            double vz = 0.0;
            if (!zz) {
              for (int cc = 0; cc < nz_; ++cc) {
                const double v = Nraw[node_index(g, sx[a], sy[bb], sz[cc]) * g.S + s];
                vz = cc == 0 ? v : vz + v;
              }
            }
            vy = bb == 0 ? vz : vy + vz;
          }
        }
        vx = a == 0 ? vy : vx + vy;
      }
    }
    N[node * g.S + s] = vx;
    cacc = cacc + vx * charge[s];          // ParticlesGridify.f90:288-293
  }
  C[node] = cacc;
}

// The same guard folds for the fluxes (ParticlesGridifyAll.f90:238-420):
// nc = 3*S components per node.
__global__ void __launch_bounds__(kT) k_fold_flux(GridDims g, const double *__restrict__ Fraw,
                                                   double *__restrict__ F) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int sx[2], sy[2], sz[2];
  bool zx, zy, zz;
  const int nx_ = fold_src(i, g.nx, g.flow[0], sx, &zx);
  const int ny_ = fold_src(j, g.ny, g.flow[1], sy, &zy);
  const int nz_ = fold_src(k, g.nz, g.flow[2], sz, &zz);
  const int nc = 3 * g.S;
  for (int q = 0; q < nc; ++q) {
    double vx = 0.0;
    if (!zx) {
      for (int a = 0; a < nx_; ++a) {
        double vy = 0.0;
        if (!zy) {
          for (int bb = 0; bb < ny_; ++bb) {
            double vz = 0.0;
            if (!zz) {
              for (int cc = 0; cc < nz_; ++cc) {
                const double v = Fraw[node_index(g, sx[a], sy[bb], sz[cc]) * nc + q];
                vz = cc == 0 ? v : vz + v;
              }
            }
            vy = bb == 0 ? vz : vy + vz;
          }
        }
        vx = a == 0 ? vy : vx + vy;
      }
    }
    F[node * nc + q] = vx;
  }
}

// ---------------------------------------------------------------------------
// CalcDistance (CalcDistance.f90:25-91), one thread per (ii, jj, kk).

__device__ __forceinline__ double inv_len(double a, double b, double c) {
  return 1.0 / sqrt(a * a + b * b + c * c);
}

__global__ void __launch_bounds__(kT) k_calc_distance(int Nx, int Ny, int Nz, double hx, double hy,
                                                       double hz, int xf, int yf, int zf,
                                                       double *__restrict__ D) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t total = (int64_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
  if (t >= total) return;
  const int ii = (int)(t % (Nx + 1)) + 1;
  const int jj = (int)((t / (Nx + 1)) % (Ny + 1)) + 1;
  const int kk = (int)(t / ((int64_t)(Nx + 1) * (Ny + 1))) + 1;
  const double di = (double)(ii - 1), dj = (double)(jj - 1), dk = (double)(kk - 1);
  double v = (ii == 1 && jj == 1 && kk == 1) ? 0.0 : inv_len(di * hx, dj * hy, dk * hz);
  const double mx = ((double)Nx - di) * hx, my = ((double)Ny - dj) * hy, mz = ((double)Nz - dk) * hz;
  if (xf) {
    if (!yf && !zf) {
      v = v + inv_len(di * hx, my, dk * hz);
      v = v + inv_len(di * hx, dj * hy, mz);
      v = v + inv_len(di * hx, my, mz);
    } else if (!yf) {
      v = v + inv_len(di * hx, my, dk * hz);
    } else if (!zf) {
      v = v + inv_len(di * hx, dj * hy, mz);
    }
  } else if (yf) {
    if (!zf) {
      v = v + inv_len(mx, dj * hy, dk * hz);
      v = v + inv_len(di * hx, dj * hy, mz);
      v = v + inv_len(mx, dj * hy, mz);
    } else {
      v = v + inv_len(mx, dj * hy, dk * hz);
    }
  } else if (zf) {
    v = v + inv_len(mx, dj * hy, dk * hz);
    v = v + inv_len(di * hx, my, dk * hz);
    v = v + inv_len(mx, my, dk * hz);
  } else {
    v = v + inv_len(mx, dj * hy, dk * hz);
    v = v + inv_len(di * hx, my, dk * hz);
    v = v + inv_len(di * hx, dj * hy, mz);
    v = v + inv_len(mx, my, dk * hz);
    v = v + inv_len(di * hx, my, mz);
    v = v + inv_len(mx, dj * hy, mz);
    v = v + inv_len(mx, my, mz);
  }
  v = 1.0 * v;
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1;
  const int pi = Nx + 1 + (ii - 1), pj = Ny + 1 + (jj - 1), pk = Nz + 1 + (kk - 1);
  const int ni = Nx + 1 - (ii - 1), nj = Ny + 1 - (jj - 1), nk = Nz + 1 - (kk - 1);
#define DAT(a, b, c) D[(int64_t)((a) - 1) + DX * ((int64_t)((b) - 1) + DY * (int64_t)((c) - 1))]
  DAT(pi, pj, pk) = v; DAT(ni, pj, pk) = v; DAT(pi, nj, pk) = v; DAT(pi, pj, nk) = v;
  DAT(ni, nj, pk) = v; DAT(pi, nj, nk) = v; DAT(ni, pj, nk) = v; DAT(ni, nj, nk) = v;
#undef DAT
}

// ---------------------------------------------------------------------------
// BCpotentials (BCpotentials.f90:51-141, one rank).  Every face node's value
// is the sequential sum, in Fortran column-major order (k, then j, then i
// fastest), of C(interior) * D(window); that is how amdflang evaluates the
// SUM of the array expression, and each lane below keeps exactly that order.
//
// Lanes of a wave take 64 consecutive face nodes along one transverse axis a
// and are skewed along the loop over that same axis: at wave step t lane l
// handles a-index t + l.  The D offset along a is (t + l) - (a0 + l), the same
// for every lane, so D is read once per wave (uniform, scalar loads) and each
// lane reads its own charge.  Each lane also accumulates R face nodes along
// the other transverse axis b, which share the charge and differ only in D.
// Steps at which a lane is outside [1, N_a] add nothing, so every lane adds
// the same products in the same order as the reference.

constexpr int kFaceR = 4;          // face nodes per lane along b
constexpr int kFaceWaves = 4;      // waves per block, along b

// x faces (lanes along y, skew on j).  Ct holds the interior charge with y
// fastest, Ct[(k-1)*Nx*Ny + (i-1)*Ny + (j-1)], so the lanes' reads coalesce.
__global__ void __launch_bounds__(64 * kFaceWaves) k_bcpot_x(GridDims g, const double *__restrict__ Ct,
                                                             const double *__restrict__ D,
                                                             double *__restrict__ U, double kfac) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform per wave
  const int face = blockIdx.z;
  const int a0 = blockIdx.x * 64 + 1;                               // jj of lane 0
  const int b0 = (blockIdx.y * kFaceWaves + wave) * kFaceR + 1;     // kk of r = 0
  if (b0 > Nz) return;
  const int La = Ny - a0 + 1 < 64 ? Ny - a0 + 1 : 64;              // active lanes
  const bool active = lane < La;
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1;
  const int di = face == 0 ? Nx + 2 : 1;                           // BCpotentials.f90:60/63
  double s[kFaceR];
#pragma unroll
  for (int r = 0; r < kFaceR; ++r) s[r] = 0.0;
  for (int k = 1; k <= Nz; ++k) {
    const double *ck = Ct + (int64_t)(k - 1) * Nx * Ny;
    for (int t = 1 - (La - 1); t <= Ny; ++t) {
      const int j = t + lane;
      const bool on = active && j >= 1 && j <= Ny;
      const int64_t ydz = (int64_t)(Ny + t - a0) * DX;              // y index Ny+1+j-jj, 0-based
      const double *drow[kFaceR];
#pragma unroll
      for (int r = 0; r < kFaceR; ++r) {
        const int bb = b0 + r <= Nz ? b0 + r : Nz;                 // clamp: unused lanes of r
        drow[r] = D + (int64_t)(di - 1) + ydz + DX * DY * (int64_t)(Nz + k - bb);
      }
      if (on) {
        const double *crow = ck + (j - 1);
        for (int i = 0; i < Nx; ++i) {
          const double c = crow[(int64_t)i * Ny];
#pragma unroll
          for (int r = 0; r < kFaceR; ++r) s[r] = s[r] + c * drow[r][i];
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int r = 0; r < kFaceR; ++r) {
    const int bb = b0 + r;
    if (bb > Nz) break;
    U[node_index(g, face == 0 ? 1 : Nx + 2, a0 + lane + 1, bb + 1)] = s[r] * kfac;
  }
}

// x faces with the charge rows staged in LDS: a block's waves share one range
// of 64 face nodes along y (lanes) and take R face nodes each along z; at
// wave step t lane l needs charge row j = t + l of plane k, so the block keeps
// rows t .. t+63 in a ring (row j in slot j % 64, stride Nx + 1 doubles to
// spread the banks) and loads one new row per step.  Each charge value is
// read from HBM once per block instead of once per lane; the summation order
// of every lane is that of k_bcpot_x.
template <int R>
__global__ void __launch_bounds__(256) k_bcpot_x_lds(GridDims g, const double *__restrict__ C,
                                                     const double *__restrict__ D,
                                                     double *__restrict__ U, double kfac) {
  extern __shared__ double ring[];
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const int RS = Nx + 1;
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform per wave
  const int face = blockIdx.z;
  const int a0 = blockIdx.x * 64 + 1;                                  // jj of lane 0
  const int b0 = (blockIdx.y * nw + wave) * R + 1;                     // kk of r = 0
  const bool wave_on = b0 <= Nz;                                       // still loads rows
  const int La = Ny - a0 + 1 < 64 ? Ny - a0 + 1 : 64;
  const bool active = wave_on && lane < La;
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1;
  const int64_t X = g.X;
  const int di = face == 0 ? Nx + 2 : 1;                               // BCpotentials.f90:60/63
  double s[R];
#pragma unroll
  for (int r = 0; r < R; ++r) s[r] = 0.0;
  for (int k = 1; k <= Nz; ++k) {
    const double *cplane = C + node_index(g, 2, 2, k + 1);            // row j at (j-1)*X
    for (int t = 1 - (La - 1); t <= Ny; ++t) {
      __syncthreads();                                                 // step t-1 done
      const int jn = t + La - 1;                                       // row first needed now
      if (jn >= 1 && jn <= Ny) {
        const double *src = cplane + (int64_t)(jn - 1) * X;
        double *dst = ring + (jn & 63) * RS;
        for (int x = threadIdx.x; x < Nx; x += blockDim.x) dst[x] = src[x];
      }
      __syncthreads();
      const int j = t + lane;
      if (active && j >= 1 && j <= Ny) {
        const int64_t ydz = (int64_t)(Ny + t - a0) * DX;               // y index Ny+1+j-jj, 0-based
        const double *drow[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int bb = b0 + r <= Nz ? b0 + r : Nz;
          drow[r] = D + (int64_t)(di - 1) + ydz + DX * DY * (int64_t)(Nz + k - bb);
        }
        const double *row = ring + (j & 63) * RS;
        // D in blocks of 8 consecutive values (one wide scalar load per row)
        int i = 0;
        for (; i + 8 <= Nx; i += 8) {
          double d[R][8], cv[8];
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < 8; ++q) d[r][q] = drow[r][i + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) cv[q] = row[i + q];
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = s[r] + cv[q] * d[r][q];
        }
        for (; i < Nx; ++i) {
          const double c = row[i];
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = s[r] + c * drow[r][i];
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int bb = b0 + r;
    if (bb > Nz) break;
    U[node_index(g, face == 0 ? 1 : Nx + 2, a0 + lane + 1, bb + 1)] = s[r] * kfac;
  }
}

// y faces (dim 1: lanes along x, b = z) and z faces (dim 2: lanes along x,
// b = y); skew on the innermost loop i, so the charge reads coalesce as is.
__global__ void __launch_bounds__(64 * kFaceWaves) k_bcpot_yz(GridDims g, int dim,
                                                              const double *__restrict__ C,
                                                              const double *__restrict__ D,
                                                              double *__restrict__ U, double kfac) {
  const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform per wave
  const int face = blockIdx.z;
  const int nb = dim == 1 ? Nz : Ny;
  const int a0 = blockIdx.x * 64 + 1;                               // ii of lane 0
  const int b0 = (blockIdx.y * kFaceWaves + wave) * kFaceR + 1;     // kk (dim 1) or jj (dim 2)
  if (b0 > nb) return;
  const int La = Nx - a0 + 1 < 64 ? Nx - a0 + 1 : 64;
  const bool active = lane < La;
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1;
  double s[kFaceR];
#pragma unroll
  for (int r = 0; r < kFaceR; ++r) s[r] = 0.0;
  for (int k = 1; k <= Nz; ++k) {
    for (int j = 1; j <= Ny; ++j) {
      // D offset of (j, k) for each r: dim 1: y fixed by the face, z by kk;
      // dim 2: y by jj, z fixed by the face (BCpotentials.f90:91/94, 122/125)
      int64_t dyz[kFaceR];
#pragma unroll
      for (int r = 0; r < kFaceR; ++r) {
        const int bb = b0 + r <= nb ? b0 + r : nb;
        int64_t yi, zi;                                             // 0-based D indices
        if (dim == 1) {
          yi = face == 0 ? (int64_t)Ny + j : (int64_t)j - 1;
          zi = (int64_t)Nz + k - bb;
        } else {
          yi = (int64_t)Ny + j - bb;
          zi = face == 0 ? (int64_t)Nz + k : (int64_t)k - 1;
        }
        dyz[r] = DX * (yi + DY * zi);
      }
      const double *crow = C + node_index(g, 2, j + 1, k + 1);
      for (int t = 1 - (La - 1); t <= Nx; ++t) {
        const int i = t + lane;
        const int64_t xi = (int64_t)Nx + t - a0;                    // x index Nx+1+i-ii, 0-based
        if (active && i >= 1 && i <= Nx) {
          const double c = crow[i - 1];
#pragma unroll
          for (int r = 0; r < kFaceR; ++r) s[r] = s[r] + c * D[xi + dyz[r]];
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int r = 0; r < kFaceR; ++r) {
    const int bb = b0 + r;
    if (bb > nb) break;
    const int ii = a0 + lane;
    const int64_t at = dim == 1 ? node_index(g, ii + 1, face == 0 ? 1 : Ny + 2, bb + 1)
                                : node_index(g, ii + 1, bb + 1, face == 0 ? 1 : Nz + 2);
    U[at] = s[r] * kfac;
  }
}

// Interior charge with y fastest, for k_bcpot_x.
__global__ void __launch_bounds__(256) k_charge_yfirst(GridDims g, const double *__restrict__ C,
                                                       double *__restrict__ Ct) {
  __shared__ double tile[32][33];
  const int k = blockIdx.z + 1;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;           // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + tx + 1, j = j0 + q + 1;
    if (i <= g.nx && j <= g.ny) tile[q][tx] = C[node_index(g, i + 1, j + 1, k + 1)];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int j = j0 + tx + 1, i = i0 + q + 1;
    if (i <= g.nx && j <= g.ny)
      Ct[(int64_t)(k - 1) * g.nx * g.ny + (int64_t)(i - 1) * g.ny + (j - 1)] = tile[tx][q];
  }
}

// ---------------------------------------------------------------------------
// CalcEpotential (CalcEpotential.f90:28-242)
//
// Jacobi iterates U_q live in three rotating buffers jb[q % 3]; all three
// start as copies of the input U (interior, flow faces, guards).  On one rank
// the reference's periodic halo of iterate q >= 1 is a copy of iterate q's
// interior at the wrapped position, and iterate 0 keeps whatever halo U came
// in with; flow faces keep their input values throughout.
//
// Pass p computes U_p = J(U_{p-1}) and, for p >= 2, the residual r_{p-1} of
// U_{p-1}, which the reference evaluates with the halo of U_{p-2} ("stale"
// halo: CalcEpotential.f90:83-92 runs before the exchange of :100-239).  The
// last block to finish a pass reduces the block sums in a fixed order and
// applies the stop test of loop iteration ss = p (:60-62), so the decision is
// taken one speculative sweep late and U_p is simply discarded on exit.

struct JacobiCoef {
  double ax, ay, az, af;
  double dx2, dy2, dz2;
};




__global__ void __launch_bounds__(kT) k_jacobi_bnorm(GridDims g, const double *__restrict__ C,
                                                      double *__restrict__ b, double binv,
                                                      double *__restrict__ partials,
                                                      JacobiState *st, double maxerr) {
  __shared__ double sh[kT / 64];
  __shared__ bool amLast;
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  double v2 = 0.0;
  if (node < g.G) {
    const double bv = -C[node] * binv;               // CalcEpotential.f90:50
    b[node] = bv;
    const int i = (int)(node % g.X) + 1;
    const int j = (int)((node / g.X) % g.Y) + 1;
    const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
    if (i >= 2 && i <= g.nx + 1 && j >= 2 && j <= g.ny + 1 && k >= 2 && k <= g.nz + 1)
      v2 = bv * bv;
  }
  const double s = block_sum(v2, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s;
    __threadfence();
    const unsigned prev = atomicAdd(&st->arrive, 1u);
    amLast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!amLast) return;
  const double tot = final_sum(partials, gridDim.x, sh);
  if (threadIdx.x == 0) {
    st->bsum = tot;
    st->rsum = tot;                                  // r_global = b_global (:55)
    st->thresh = maxerr * sqrt(tot);
    st->arrive = 0;
    // first stop test, ss = 1 (:60-62); the decision goes to slot 0 (launch 0)
    const bool stop = st->maxit >= 1 && sqrt(tot) <= st->thresh;
    st->slot[0].done = stop ? 1 : 0;
    st->slot[0].result = 0;
    if (stop) {
      st->done = 1;
      st->result = 0;
    }
  }
}

// Sum of the previous launch's per-block partials, by every block in the same
// fixed order (no atomics, no fences: the kernel boundary orders the data).
__device__ __forceinline__ double sum_partials(const double *__restrict__ P, int n, double *sh) {
  double v = 0.0;
  const int nt = (int)(blockDim.x * blockDim.y * blockDim.z);
  for (int q = lin_tid(); q < n; q += nt) v = v + P[q];
  return block_sum(v, sh);
}

// Uniform geometry of the pass kernel, kept small (32-bit, no box data) so
// it stays in scalar registers.
struct PassGeo {
  int32_t nx, ny, nz;
  uint32_t X, XY;        // node strides of j and k
  int32_t px, py, pz;    // periodic flags
  int32_t ntx, TP, W;    // tile schedule (tile width = blockDim.x)
};

__global__ void __launch_bounds__(kStencilT, 8) k_jacobi_pass(PassGeo pg, JacobiCoef jc, double *b0, double *b1,
                                                     double *b2, const double *__restrict__ b,
                                                     const double *__restrict__ Pin,
                                                     double *__restrict__ Pout, JacobiState *st,
                                                     int32_t p, unsigned long long *ks) {
  __shared__ double sh[kStencilT / 64];
  __shared__ int decide;
  kstamp_begin(ks);
  const StopSlot prev = st->slot[(p - 1) & 1];
  const int t0 = lin_tid();
  // decision of this launch: carried over, or the stop test of loop
  // iteration ss = p-1 on the residual of U_{p-2} (partials of launch p-1)
  StopSlot cur = prev;
  if (!prev.done) {
    if (p >= 3) {
      const double S = sum_partials(Pin, (int)gridDim.x, sh);
      if (t0 == 0) decide = (p - 1 <= st->maxit && sqrt(S) <= st->thresh) ? 1 : 0;
      __syncthreads();
      if (decide) {
        cur.done = 1;
        cur.result = p - 2;
      }
    }
    if (!cur.done && p > st->last) {      // all 2*maxit+1 sweeps done (:58)
      cur.done = 1;
      cur.result = st->last;
    }
  }
  if (blockIdx.x == 0 && t0 == 0) st->slot[p & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }

  const int r = p % 3;
  const double *curb = r == 1 ? b0 : (r == 2 ? b1 : b2);           // U_{p-1} in bufs[(p-1)%3]
  double *out = r == 0 ? b0 : (r == 1 ? b1 : b2);                  // U_p in bufs[p%3]
  const double *old = p >= 2 ? (r == 2 ? b0 : (r == 0 ? b1 : b2)) : nullptr;   // U_{p-2}
  const bool wrap_cur = (p - 1) >= 1;              // halo of U_{p-1} = wrapped interior
  const bool wrap_old = (p - 2) >= 1;
  const bool px = pg.px, py = pg.py, pz = pg.pz;
  const uint32_t X = pg.X, XY = pg.XY;
  const int nx = pg.nx, ny = pg.ny, nz = pg.nz;
  // plane index (0-based) of 1-based plane kk of the iterate, honouring the halo rule
  auto zplane = [&](int kk, bool wrap) -> uint32_t {
    if (pz && wrap) {
      if (kk == 1) return (uint32_t)nz;
      if (kk == nz + 2) return 1u;
    }
    return (uint32_t)(kk - 1);
  };
  double r2 = 0.0;
  // 2.5D march: each thread owns an (i, j) column of its XCD's z-slab and
  // keeps the z-neighbours of U_{p-1} in registers
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  for (int t = w; z1 > z0 && t < pg.TP; t += pg.W) {
    const int tx = t % pg.ntx, ty = t / pg.ntx;
    const int i = tx * (int)blockDim.x + (int)threadIdx.x + 2;
    const int j = ty * (int)blockDim.y + (int)threadIdx.y + 2;
    if (i > nx + 1 || j > ny + 1) continue;
    const uint32_t col = (uint32_t)(i - 1) + X * (uint32_t)(j - 1);
    const uint32_t oe = (px && wrap_cur && i == nx + 1) ? 0u - (uint32_t)(nx - 1) : 1u;
    const uint32_t ow = (px && wrap_cur && i == 2) ? (uint32_t)(nx - 1) : 0u - 1u;
    const uint32_t on = (py && wrap_cur && j == ny + 1) ? 0u - (uint32_t)(ny - 1) * X : X;
    const uint32_t os = (py && wrap_cur && j == 2) ? (uint32_t)(ny - 1) * X : 0u - X;
    const bool edge_xy = (px && (i == nx + 1 || i == 2)) || (py && (j == ny + 1 || j == 2));
    int k = z0 + 2;
    double zm = curb[col + XY * zplane(k - 1, wrap_cur)];
    double zc = curb[col + XY * (uint32_t)(k - 1)];
    double zp = curb[col + XY * zplane(k + 1, wrap_cur)];
#pragma unroll 2
    for (; k <= z1 + 1; ++k) {
      const uint32_t c = col + XY * (uint32_t)(k - 1);
      const double ue = curb[c + oe], uw = curb[c + ow], un = curb[c + on], us = curb[c + os];
      const double ut = zp, ub = zm;
      const double bc = b[c];
      // CalcEpotential.f90:74-80
      out[c] = jc.ax * (ue + uw) + jc.ay * (un + us) + jc.az * (ut + ub) - jc.af * bc;
      if (p >= 2) {
        // residual of U_{p-1} with the halo of U_{p-2}: periodic guard
        // neighbours come from `old` (wrapped, or the stored input halo at p-2 = 0)
        double re = ue, rw = uw, rn = un, rs = us, rt = ut, rb = ub;
        if (edge_xy || (pz && (k == nz + 1 || k == 2))) {
          if (px && i == nx + 1) re = wrap_old ? old[c - (uint32_t)(nx - 1)] : old[c + 1];
          if (px && i == 2) rw = wrap_old ? old[c + (uint32_t)(nx - 1)] : old[c - 1];
          if (py && j == ny + 1) rn = wrap_old ? old[c - (uint32_t)(ny - 1) * X] : old[c + X];
          if (py && j == 2) rs = wrap_old ? old[c + (uint32_t)(ny - 1) * X] : old[c - X];
          if (pz && k == nz + 1) rt = wrap_old ? old[c - (uint32_t)(nz - 1) * XY] : old[c + XY];
          if (pz && k == 2) rb = wrap_old ? old[c + (uint32_t)(nz - 1) * XY] : old[c - XY];
        }
        const double u2 = 2.0 * zc;
        // CalcEpotential.f90:83-92
        const double res = bc - ((re + rw - u2) / jc.dx2 + (rn + rs - u2) / jc.dy2 +
                                 (rt + rb - u2) / jc.dz2);
        r2 = r2 + res * res;
      }
      zm = zc;
      zc = zp;
      if (k + 1 <= z1 + 1) zp = curb[col + XY * zplane(k + 2, wrap_cur)];
    }
  }
  const double sblk = block_sum(r2, sh);
  if (t0 == 0) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// The same pass with two cells per lane, for even nx: a wave owns 128 cells
// of a row, each lane's pair moves with one 16-byte access (the iterate
// buffers and b start one element past a 16-byte boundary, so that the pairs
// (i, i+1), i even, are aligned), and the z-march is software-pipelined: four
// planes per pass with the z-window registers rotating by name and each
// plane's loads issued two planes ahead.  Each lane reads its pair's two
// x-neighbours itself (8-byte loads from the lines its wave already touches):
// no cross-lane exchange, which the compiler may not keep convergent across
// the lane-dependent branches of the boundary loads.  The stale-halo values
// of U_{p-2} on the y and z boundaries are read where they are used (rows 2
// and ny+1, planes 2 and nz+1: whole waves, rarely); those of the x boundary
// travel with the plane's loads.  Every cell's arithmetic is k_jacobi_pass's,
// so the iterates are the same.
// A 16-byte store marked non-temporal: U_p is not read again in this pass,
// and the 400 MB a pass moves at 256^3 do not fit the last-level cache anyway,
// so the hint keeps the stream from displacing lines still to be read
// (measured: 93 us per pass at 256^3 against 102 us with plain stores).
__device__ __forceinline__ void st2_stream(double *p, double a, double b) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v t = {a, b};
  __builtin_nontemporal_store(t, reinterpret_cast<d2v *>(p));
}

template <bool PX>
__global__ void __launch_bounds__(kStencilT, 4) k_jacobi_pass2(PassGeo pg, JacobiCoef jc,
                                                       double *b0, double *b1, double *b2,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ Pin,
                                                       double *__restrict__ Pout, JacobiState *st,
                                                       int32_t p, unsigned long long *ks) {
  __shared__ double sh[kStencilT / 64];
  __shared__ int decide;
  kstamp_begin(ks);
  const StopSlot prev = st->slot[(p - 1) & 1];
  const int t0 = lin_tid();
  StopSlot cur = prev;
  if (!prev.done) {
    if (p >= 3) {
      const double S = sum_partials(Pin, (int)gridDim.x, sh);
      if (t0 == 0) decide = (p - 1 <= st->maxit && sqrt(S) <= st->thresh) ? 1 : 0;
      __syncthreads();
      if (decide) {
        cur.done = 1;
        cur.result = p - 2;
      }
    }
    if (!cur.done && p > st->last) {      // all 2*maxit+1 sweeps done (:58)
      cur.done = 1;
      cur.result = st->last;
    }
  }
  if (blockIdx.x == 0 && t0 == 0) st->slot[p & 1] = cur;
  if (cur.done) {
    kstamp_end(ks);
    return;
  }

  const int r = p % 3;
  const double *curb = r == 1 ? b0 : (r == 2 ? b1 : b2);           // U_{p-1} in bufs[(p-1)%3]
  double *out = r == 0 ? b0 : (r == 1 ? b1 : b2);                  // U_p in bufs[p%3]
  const double *old = p >= 2 ? (r == 2 ? b0 : (r == 0 ? b1 : b2)) : nullptr;   // U_{p-2}
  const bool wrap_cur = (p - 1) >= 1;              // halo of U_{p-1} = wrapped interior
  const bool wrap_old = (p - 2) >= 1;
  const bool res = p >= 2;
  const bool py = pg.py, pz = pg.pz;
  const uint32_t X = pg.X, XY = pg.XY;
  const int nx = pg.nx, ny = pg.ny, nz = pg.nz;
  const int lane = (int)threadIdx.x;
  // the index is formed in 32 bits first: offsets below are unsigned and wrap
  auto ld2 = [](const double *base, uint32_t idx) {
    return *reinterpret_cast<const double2 *>(base + idx);
  };
  auto zplane = [&](int kk, bool wrap) -> uint32_t {
    if (pz && wrap) {
      if (kk == 1) return (uint32_t)nz;
      if (kk == nz + 2) return 1u;
    }
    return (uint32_t)(kk - 1);
  };
  // offsets from a cell to its stale-halo neighbours in U_{p-2}
  const uint32_t rwo = wrap_old ? (uint32_t)(nx - 1) : 0u - 1u;            // west of i = 2
  const uint32_t reo = wrap_old ? 0u - (uint32_t)(nx - 1) : 1u;            // east of i = nx+1
  const uint32_t rno = wrap_old ? 0u - (uint32_t)(ny - 1) * X : X;         // north of j = ny+1
  const uint32_t rso = wrap_old ? (uint32_t)(ny - 1) * X : 0u - X;         // south of j = 2
  const uint32_t rto = wrap_old ? 0u - (uint32_t)(nz - 1) * XY : XY;       // top of k = nz+1
  const uint32_t rbo = wrap_old ? (uint32_t)(nz - 1) * XY : 0u - XY;       // bottom of k = 2
  double r2 = 0.0;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int z0 = (xcd * nz) >> 3, z1 = ((xcd + 1) * nz) >> 3;
  for (int t = w; z1 > z0 && t < pg.TP; t += pg.W) {
    const int tx = t % pg.ntx, ty = t / pg.ntx;
    const int j = ty * (int)blockDim.y + (int)threadIdx.y + 2;
    if (j > ny + 1) continue;                                // the whole wave
    const int ia = tx * 128 + 2 * lane + 2;                  // the pair (ia, ia+1), 1-based
    const bool live = ia <= nx;
    const int i0 = live ? ia : nx;                           // past the row end: the last pair again
    const uint32_t col = (uint32_t)(i0 - 1) + X * (uint32_t)(j - 1);
    const bool hasf = i0 == 2, hasl = i0 == nx;              // the pair holds i = 2 / i = nx+1
    const uint32_t ow = (PX && wrap_cur && hasf) ? (uint32_t)(nx - 1) : 0u - 1u;   // from i0
    const uint32_t oe = (PX && wrap_cur && hasl) ? 0u - (uint32_t)(nx - 1) : 1u;   // from i0+1
    const uint32_t on = (py && wrap_cur && j == ny + 1) ? 0u - (uint32_t)(ny - 1) * X : X;
    const uint32_t os = (py && wrap_cur && j == 2) ? (uint32_t)(ny - 1) * X : 0u - X;
    const bool yN = py && j == ny + 1, yS = py && j == 2;
    struct PIn {
      double2 n, s, bc;
      double w, e, sw, se;
    };
    auto zval = [&](int kk) -> double2 { return ld2(curb, col + XY * zplane(kk, wrap_cur)); };
    auto plane_in = [&](int kk) -> PIn {
      const uint32_t c = col + XY * (uint32_t)(kk - 1);
      PIn a;
      a.sw = 0.0;
      a.se = 0.0;
      if (PX && res && hasf) a.sw = old[c + rwo];
      if (PX && res && hasl) a.se = old[c + 1 + reo];
      a.w = curb[c + ow];
      a.e = curb[c + 1 + oe];
      a.n = ld2(curb, c + on);
      a.s = ld2(curb, c + os);
      a.bc = ld2(b, c);
      return a;
    };
    // plane kk from its loads and the z-window (zb, zc, zt) = planes kk-1, kk, kk+1
    auto do_plane = [&](int kk, const PIn &a, const double2 &zb, const double2 &zc,
                        const double2 &zt) {
      const uint32_t c = col + XY * (uint32_t)(kk - 1);
      const double wv = a.w, ev = a.e;
      const double uc[2] = {zc.x, zc.y};
      const double un[2] = {a.n.x, a.n.y}, us[2] = {a.s.x, a.s.y};
      const double ut[2] = {zt.x, zt.y}, ub[2] = {zb.x, zb.y};
      const double bc[2] = {a.bc.x, a.bc.y};
      double v[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double uw = q == 0 ? wv : uc[0];
        const double ue = q == 1 ? ev : uc[1];
        // CalcEpotential.f90:74-80
        v[q] = jc.ax * (ue + uw) + jc.ay * (un[q] + us[q]) + jc.az * (ut[q] + ub[q]) -
               jc.af * bc[q];
      }
      if (live) st2_stream(out + c, v[0], v[1]);
      if (res) {
        // residual of U_{p-1} with the halo of U_{p-2} across periodic boundaries
        double rn[2] = {un[0], un[1]}, rs[2] = {us[0], us[1]};
        double rt[2] = {ut[0], ut[1]}, rb[2] = {ub[0], ub[1]};
        if (yN || yS || (pz && (kk == nz + 1 || kk == 2))) {
          if (yN) {
            const double2 q2 = ld2(old, c + rno);
            rn[0] = q2.x;
            rn[1] = q2.y;
          }
          if (yS) {
            const double2 q2 = ld2(old, c + rso);
            rs[0] = q2.x;
            rs[1] = q2.y;
          }
          if (pz && kk == nz + 1) {
            const double2 q2 = ld2(old, c + rto);
            rt[0] = q2.x;
            rt[1] = q2.y;
          }
          if (pz && kk == 2) {
            const double2 q2 = ld2(old, c + rbo);
            rb[0] = q2.x;
            rb[1] = q2.y;
          }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const double uw = q == 0 ? wv : uc[0];
          const double ue = q == 1 ? ev : uc[1];
          const double re = (PX && hasl && q == 1) ? a.se : ue;
          const double rw = (PX && hasf && q == 0) ? a.sw : uw;
          const double u2 = 2.0 * uc[q];
          // CalcEpotential.f90:83-92
          const double rr = bc[q] - ((re + rw - u2) / jc.dx2 + (rn[q] + rs[q] - u2) / jc.dy2 +
                                     (rt[q] + rb[q] - u2) / jc.dz2);
          if (live) r2 = r2 + rr * rr;
        }
      }
    };
    const int ka = z0 + 2, kb = z1 + 2;                      // planes [ka, kb), 1-based
    auto zl = [&](int kk) { return zval(kk < kb ? kk : kb); };
    auto pl = [&](int kk) { return plane_in(kk < kb ? kk : kb - 1); };
    double2 w0 = zl(ka - 1), w1 = zl(ka), w2 = zl(ka + 1), w3 = zl(ka + 2);
    PIn A = pl(ka), B = pl(ka + 1);
    for (int k = ka; k < kb; k += 4) {
      do_plane(k, A, w0, w1, w2);
      if (k + 1 >= kb) break;
      w0 = zl(k + 3);
      A = pl(k + 2);
      do_plane(k + 1, B, w1, w2, w3);
      if (k + 2 >= kb) break;
      w1 = zl(k + 4);
      B = pl(k + 3);
      do_plane(k + 2, A, w2, w3, w0);
      if (k + 3 >= kb) break;
      w2 = zl(k + 5);
      A = pl(k + 4);
      do_plane(k + 3, B, w3, w0, w1);
      w3 = zl(k + 6);
      B = pl(k + 5);
    }
  }
  const double sblk = block_sum(r2, sh);
  if (t0 == 0) Pout[blockIdx.x] = sblk;
  kstamp_end(ks);
}

// After q >= 1 sweeps: U(node) = U_q at the composite periodic image of the
// node (z, then y, then x copies of whole planes, CalcEpotential.f90:100-239);
// nodes whose image lies on a flow face keep the input value found there.
__global__ void __launch_bounds__(kT) k_jacobi_finish(GridDims g, double *b0, double *b1,
                                                       double *b2, double *U, int64_t q) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  if (q == 0) return;
  const double *src_buf = q % 3 == 0 ? b0 : (q % 3 == 1 ? b1 : b2);
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int si = i, sj = j, sk = k;
  if (!g.flow[0]) si = i == 1 ? g.nx + 1 : (i == g.nx + 2 ? 2 : i);
  if (!g.flow[1]) sj = j == 1 ? g.ny + 1 : (j == g.ny + 2 ? 2 : j);
  if (!g.flow[2]) sk = k == 1 ? g.nz + 1 : (k == g.nz + 2 ? 2 : k);
  const bool interior = si >= 2 && si <= g.nx + 1 && sj >= 2 && sj <= g.ny + 1 && sk >= 2 &&
                        sk <= g.nz + 1;
  const int64_t src = node_index(g, si, sj, sk);
  if (interior) U[node] = src_buf[src];
  else if (src != node) U[node] = src_buf[src];    // flow-face values are the same in all buffers
}

// ---------------------------------------------------------------------------
// CalcEfield (CalcEfield.f90:27-228) + the on-rank halo copies: every node
// takes the central difference at its composite periodic image, or zero when
// that image lies in a flow guard (never filled on one rank).

__global__ void __launch_bounds__(kT) k_efield(GridDims g, const double *__restrict__ U,
                                                double *__restrict__ E) {
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
  E[3 * node + 0] = ex;
  E[3 * node + 1] = ey;
  E[3 * node + 2] = ez;
}

// BCfields.f90:31-90 on one rank: zero the guard planes of flow faces.
__global__ void __launch_bounds__(kT) k_bc_fields(GridDims g, double *__restrict__ E) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  const bool z = (g.flow[2] && (k == 1 || k == g.Z)) || (g.flow[1] && (j == 1 || j == g.Y)) ||
                 (g.flow[0] && (i == 1 || i == g.X));
  if (z) {
    E[3 * node + 0] = 0.0;
    E[3 * node + 1] = 0.0;
    E[3 * node + 2] = 0.0;
  }
}

// Order-free FP64 deposit (PICARD_DEPOSIT_FIXED_POINT): each particle's
// weights F1..F8 of ParticlesGridify.f90:81-88 in FP64 as the reference forms
// them, rounded to int64 fixed point (2^-48 units) and added with integer
// atomics, so N does not depend on the order of the additions; within
// 2^-49 per weight of the sequential sum.
constexpr double kFix48 = 281474976710656.0;          // 2^48
constexpr double kFix48Inv = 1.0 / 281474976710656.0;

__global__ void __launch_bounds__(kT) k_dep_fix64(ParticlesF64 p, int64_t n, GridDims g,
                                                  unsigned long long *__restrict__ Nfix,
                                                  int32_t *__restrict__ errors) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= n) return;
  const CicLoc L = cic_locate(g, p.r[0][q], p.r[1][q], p.r[2][q]);
  const int s0 = p.sp[q] - 1;
  if (!cic_in_bounds(g, L) || s0 < 0 || s0 >= g.S) {
    atomicAdd(errors, 1);
    return;
  }
  const int64_t base = node_index(g, L.idx[0], L.idx[1], L.idx[2]);
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const double fx = (o & 1) ? L.rest[0] : L.drest[0];
    const double fy = (o & 2) ? L.rest[1] : L.drest[1];
    const double fz = (o & 4) ? L.rest[2] : L.drest[2];
    const double F = fx * fy * fz / g.dV;
    const int64_t node = base + (o & 1) + (int64_t)g.X * ((o >> 1) & 1) +
                         (int64_t)g.X * g.Y * (o >> 2);
    atomicAdd(&Nfix[node * g.S + s0], (unsigned long long)llrint(F * kFix48));
  }
}

// Integer guard folds of the fixed-point sums, then N = W 2^-48 n2p and
// C = sum_s N charge (ParticlesGridify.f90:129-293).
__global__ void __launch_bounds__(kT) k_fold_fix64(GridDims g,
                                                   const unsigned long long *__restrict__ Nfix,
                                                   double *__restrict__ N, double *__restrict__ C,
                                                   const double *consts) {
  const int64_t node = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (node >= g.G) return;
  const int i = (int)(node % g.X) + 1;
  const int j = (int)((node / g.X) % g.Y) + 1;
  const int k = (int)(node / ((int64_t)g.X * g.Y)) + 1;
  int sx[2], sy[2], sz[2];
  bool zx, zy, zz;
  const int nxs = fold_src(i, g.nx, g.flow[0], sx, &zx);
  const int nys = fold_src(j, g.ny, g.flow[1], sy, &zy);
  const int nzs = fold_src(k, g.nz, g.flow[2], sz, &zz);
  double cacc = 0.0;
  for (int s = 0; s < g.S; ++s) {
    unsigned long long w = 0;
    if (!zx && !zy && !zz)
      for (int a = 0; a < nxs; ++a)
        for (int b = 0; b < nys; ++b)
          for (int c = 0; c < nzs; ++c) w += Nfix[node_index(g, sx[a], sy[b], sz[c]) * g.S + s];
    const double v = (double)w * kFix48Inv * consts[s];
    N[node * g.S + s] = v;
    cacc = cacc + v * consts[8 + s];
  }
  C[node] = cacc;
}

__global__ void k_zero_u64(unsigned long long *p, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q < n) p[q] = 0ull;
}

__global__ void k_kstamp_reset(unsigned long long *ks, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;        // word
  if (q < n * picard_ctx::kStampWords)
    ks[q] = (q % picard_ctx::kStampWords) < kStampWays ? ~0ull : 0ull;
}

}  // namespace

int kstamp_reset(picard_ctx *c, int n) {
  if (n <= 0) return PICARD_OK;
  static_assert(picard_ctx::kStampWords == 2 * kStampWays, "stamp layout");
  k_kstamp_reset<<<(n * picard_ctx::kStampWords + 255) / 256, 256, 0, c->stream>>>(c->kst, n);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

// The stream has been synchronized: the stamps of the n sampled launches are
// final (a launch that did not run leaves its pair reset and counts nothing).
int kstamp_collect(picard_ctx *c, int n) {
  if (n <= 0) return PICARD_OK;
  const int W = picard_ctx::kStampWords;
  c->kst_host.resize((size_t)W * n);
  PICARD_HIP(hipMemcpy(c->kst_host.data(), c->kst, sizeof(unsigned long long) * W * n,
                       hipMemcpyDeviceToHost));
  for (int q = 0; q < n; ++q) {
    unsigned long long a = ~0ull, b = 0ull;
    for (int w = 0; w < kStampWays; ++w) {
      a = std::min(a, c->kst_host[(size_t)W * q + w]);
      b = std::max(b, c->kst_host[(size_t)W * q + kStampWays + w]);
    }
    if (b > a) {
      c->tk_ms += (double)(b - a) * 1e-5;                // 100 MHz ticks
      c->tk_launches += 1;
    }
  }
  return PICARD_OK;
}

// ---------------------------------------------------------------------------

static int launch_gridify_fixed(picard_ctx *c) {
  const GridDims &g = c->gd;
  const int64_t n = c->n;
  PICARD_HIP(hipMemsetAsync(c->s.counters, 0, sizeof(int32_t), c->stream));
  k_zero_u64<<<grid_blocks((int64_t)g.S * g.G, kT), kT, 0, c->stream>>>(c->Nfix, (int64_t)g.S * g.G);
  if (n > 0)
    k_dep_fix64<<<grid_blocks(n, kT), kT, 0, c->stream>>>(c->p, n, g, c->Nfix, c->s.counters);
  k_fold_fix64<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->Nfix, c->N, c->C, c->consts);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipMemcpyAsync(c->host_counters, c->s.counters, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (c->host_counters[0] != 0) {
    set_error("ParticlesGridify: " + std::to_string(c->host_counters[0]) +
              " particle(s) outside the grid arrays were not deposited");
    return PICARD_ERR_RANGE;
  }
  return PICARD_OK;
}

int launch_gridify(picard_ctx *c, bool flux) {
  if (!flux && c->order_free) return launch_gridify_fixed(c);
  const GridDims &g = c->gd;
  const int64_t n = c->n;
  const int64_t nb = (int64_t)g.S * c->K;
  PICARD_HIP(hipMemsetAsync(c->s.hist, 0, sizeof(int32_t) * (nb + 1), c->stream));
  PICARD_HIP(hipMemsetAsync(c->s.cursor, 0, sizeof(int32_t) * nb, c->stream));
  PICARD_HIP(hipMemsetAsync(c->s.counters, 0, sizeof(int32_t), c->stream));
  if (n > 0) {
    k_dep_key<<<grid_blocks(n, kT), kT, 0, c->stream>>>(c->p, n, g, c->s.key, c->s.hist,
                                                        c->s.counters);
    PICARD_HIP(hipGetLastError());
  }
  PICARD_TRY(scan_exclusive_i32(c, c->s.hist, c->s.hist, nb));
  if (n > 0) {
    k_dep_scatter<<<grid_blocks(n, kT), kT, 0, c->stream>>>(n, c->s.key, c->s.hist, c->s.cursor,
                                                            c->s.perm);
    PICARD_HIP(hipGetLastError());
    k_dep_bucket_sort<<<grid_blocks(nb, kT), kT, 0, c->stream>>>(nb, c->s.hist, c->s.perm);
    PICARD_HIP(hipGetLastError());
  }
  if (flux) {
    PICARD_TRY(ensure_flux(c));
    k_dep_gather<true><<<grid_blocks(g.G * g.S, kT), kT, 0, c->stream>>>(
        c->p, g, c->consts, c->s.hist, c->s.perm, c->Nraw, c->Fraw);
    PICARD_HIP(hipGetLastError());
    k_fold_flux<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->Fraw, c->F);
  } else {
    k_dep_gather<false><<<grid_blocks(g.G * g.S, kT), kT, 0, c->stream>>>(
        c->p, g, c->consts, c->s.hist, c->s.perm, c->Nraw, nullptr);
  }
  PICARD_HIP(hipGetLastError());
  k_fold_charge<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->Nraw, c->N, c->C,
                                                            c->consts + 8);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipMemcpyAsync(c->host_counters, c->s.counters, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (c->host_counters[0] != 0) {
    set_error("ParticlesGridify: " + std::to_string(c->host_counters[0]) +
              " particle(s) outside the grid arrays were not deposited");
    return PICARD_ERR_RANGE;
  }
  return PICARD_OK;
}

int launch_calc_distance(picard_ctx *c) {
  const GridDims &g = c->gd;
  const int64_t total = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int64_t DN = (2 * (int64_t)g.nx + 1) * (2 * (int64_t)g.ny + 1) * (2 * (int64_t)g.nz + 1);
  PICARD_HIP(hipMemsetAsync(c->D, 0, sizeof(double) * DN, c->stream));
  k_calc_distance<<<grid_blocks(total, kT), kT, 0, c->stream>>>(
      g.nx, g.ny, g.nz, g.dxyz[0], g.dxyz[1], g.dxyz[2], g.flow[0], g.flow[1], g.flow[2], c->D);
  PICARD_HIP(hipGetLastError());
  c->have_D = true;
  c->fft_ready = false;
  return PICARD_OK;
}

int launch_bc_potentials(picard_ctx *c) {
  const GridDims &g = c->gd;
  if (!c->have_D) {
    set_error("BCpotentials: the distance table is not set (call CalcDistance first)");
    return PICARD_ERR_STATE;
  }
  PICARD_HIP(hipMemsetAsync(c->U, 0, sizeof(double) * g.G, c->stream));   // :36
  const double tau = 6.2831853071795864769252867665590;
  const double eps0 = 8.854187817e-12;
  const double kfac = 1.0 / (2.0 * tau * eps0);
  const int per_block = kFaceR * kFaceWaves;
  constexpr int kLdsR = 2, kLdsWaves = 4;
  const size_t ring_bytes = sizeof(double) * 64 * (size_t)(g.nx + 1);
  bool lds_ok = g.flow[0] && ring_bytes <= 160 * 1024;
  if (lds_ok && !c->bcpot_lds_set) {
    lds_ok = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bcpot_x_lds<kLdsR>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ring_bytes) == hipSuccess;
    (void)hipGetLastError();
    c->bcpot_lds_set = lds_ok;
  }
  if (lds_ok) {
    const int per_block = kLdsR * kLdsWaves;
    k_bcpot_x_lds<kLdsR><<<dim3((g.ny + 63) / 64, (g.nz + per_block - 1) / per_block, 2),
                            64 * kLdsWaves, ring_bytes, c->stream>>>(g, c->C, c->D, c->U, kfac);
    PICARD_HIP(hipGetLastError());
  } else if (g.flow[0]) {
    // the third Jacobi buffer is free until CalcEpotential copies U into it
    double *Ct = c->jb[2];
    k_charge_yfirst<<<dim3((g.nx + 31) / 32, (g.ny + 31) / 32, g.nz), 256, 0, c->stream>>>(g, c->C,
                                                                                         Ct);
    k_bcpot_x<<<dim3((g.ny + 63) / 64, (g.nz + per_block - 1) / per_block, 2), 64 * kFaceWaves, 0,
                c->stream>>>(g, Ct, c->D, c->U, kfac);
    PICARD_HIP(hipGetLastError());
  }
  for (int dim = 1; dim < 3; ++dim) {
    if (!g.flow[dim]) continue;
    const int nb = dim == 1 ? g.nz : g.ny;
    k_bcpot_yz<<<dim3((g.nx + 63) / 64, (nb + per_block - 1) / per_block, 2), 64 * kFaceWaves, 0,
                 c->stream>>>(g, dim, c->C, c->D, c->U, kfac);
    PICARD_HIP(hipGetLastError());
  }
  return PICARD_OK;
}

// b = -C/(dV eps0), its interior norm, the threshold and the stop test of
// loop iteration ss = 1 (CalcEpotential.f90:50-62).
int launch_bnorm(picard_ctx *c, double maxerr) {
  const GridDims &g = c->gd;
  const double binv = 1.0 / (g.dV * 8.854187817e-12);
  k_jacobi_bnorm<<<grid_blocks(g.G, kT), kT, 0, c->stream>>>(g, c->C, c->b, binv, c->s.partials,
                                                             c->js, maxerr);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int launch_calc_epotential(picard_ctx *c, double maxerr, int32_t maxit, int64_t *sweeps) {
  const GridDims &g = c->gd;
  const double dx = g.dxyz[0], dy = g.dxyz[1], dz = g.dxyz[2];
  JacobiCoef jc;
  const double den = 1.0 / ((dx * dy) * (dx * dy) + (dx * dz) * (dx * dz) + (dy * dz) * (dy * dz));
  jc.ax = 0.5 * den * ((dy * dz) * (dy * dz));            // CalcEpotential.f90:44-47
  jc.ay = 0.5 * den * ((dx * dz) * (dx * dz));
  jc.az = 0.5 * den * ((dx * dy) * (dx * dy));
  jc.af = 0.5 * den * ((dx * dy * dz) * (dx * dy * dz));
  jc.dx2 = dx * dx;
  jc.dy2 = dy * dy;
  jc.dz2 = dz * dz;
  JacobiState init{};
  init.pass = 1;
  init.maxit = maxit;
  init.last = 2 * (int64_t)maxit + 1;
  PICARD_HIP(hipMemcpyAsync(c->js, &init, sizeof(JacobiState), hipMemcpyHostToDevice, c->stream));
  for (int q = 0; q < 3; ++q)
    PICARD_HIP(hipMemcpyAsync(c->jb[q], c->U, sizeof(double) * g.G, hipMemcpyDeviceToDevice,
                              c->stream));
  const int nbn = grid_blocks(g.G, kT);
  PICARD_TRY(launch_bnorm(c, maxerr));

  // even nx: two cells per lane, tiles of 128 x 4 cells
  const bool two = g.nx % 2 == 0;
  TileSched ts = make_sched(g.nx, g.ny, g.nz, kStencilT);
  if (two) {
    ts.bx = 64;
    ts.by = kStencilT / 64;
    ts.ntx = (g.nx + 127) / 128;
    ts.nty = (g.ny + ts.by - 1) / ts.by;
    ts.TP = ts.ntx * ts.nty;
    ts.W = ts.TP < 256 ? (ts.TP < 4 ? 4 : ts.TP) : 256;
    ts.B = 8 * ts.W;
    if (c->pass2_wmax <= 0) {
      int per_cu = 0, cus = 0;
      const void *kf = g.flow[0] ? reinterpret_cast<const void *>(k_jacobi_pass2<false>)
                                 : reinterpret_cast<const void *>(k_jacobi_pass2<true>);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, kStencilT, 0) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
          per_cu <= 0 || cus < 8)
        c->pass2_wmax = 128;
      else
        c->pass2_wmax = (cus / 8) * per_cu;
    }
    if (ts.W > c->pass2_wmax) {
      ts.W = c->pass2_wmax < 4 ? 4 : c->pass2_wmax;
      ts.B = 8 * ts.W;
    }
  } else if (c->pass_wmax <= 0) {
    // blocks per XCD resident at once: the XCD slab schedule in one round
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_pass, kStencilT, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        per_cu <= 0 || cus < 8)
      c->pass_wmax = 256;
    else
      c->pass_wmax = (cus / 8) * per_cu;
  }
  if (!two && ts.W > c->pass_wmax) {
    ts.W = c->pass_wmax < 4 ? 4 : c->pass_wmax;
    ts.B = 8 * ts.W;
  }
  PassGeo pg;
  pg.nx = g.nx; pg.ny = g.ny; pg.nz = g.nz;
  pg.X = (uint32_t)g.X;
  pg.XY = (uint32_t)g.X * (uint32_t)g.Y;
  pg.px = !g.flow[0]; pg.py = !g.flow[1]; pg.pz = !g.flow[2];
  pg.ntx = ts.ntx; pg.TP = ts.TP; pg.W = ts.W;
  const int64_t last = init.last;
  double *P[2] = {c->s.partials, c->s.partials + c->s.partials_len};
  int64_t p = 1;                  // next pass (launch number)
  // first batch from the previous solve's count (cold starts are alike from
  // step to step), then short batches: few no-op passes after convergence
  int64_t batch = c->last_sweeps > 64 ? (c->last_sweeps * 15) / 16 : 32;
  const int64_t kMaxBatch = int64_t(1) << 16;
  if (batch > kMaxBatch) batch = kMaxBatch;
  StopSlot fin{};
  for (;;) {
    const int64_t todo = (last + 1 - (p - 1)) < batch ? (last + 1 - (p - 1)) : batch;
    // timing: one pass in kTimeEvery, with dispatch-packet events around it
    // and the kernel's own clock stamps (events around every pass would slow
    // the timed solve itself)
    int64_t timed = (todo + kTimeEvery / 2 - 1) / kTimeEvery;         // passes q = 8, 24, ...
    if (timed > picard_ctx::kStamps) timed = picard_ctx::kStamps;
    if (c->timing) {
      while ((int64_t)c->tev.size() < 2 * timed) {
        hipEvent_t e;
        PICARD_HIP(hipEventCreate(&e));
        c->tev.push_back(e);
      }
      PICARD_TRY(kstamp_reset(c, (int)timed));
    }
    for (int64_t q = 0; q < todo; ++q, ++p) {
      const bool stamp = c->timing && q % kTimeEvery == kTimeEvery / 2 && q / kTimeEvery < timed;
      unsigned long long *ks = stamp ? c->kst + picard_ctx::kStampWords * (q / kTimeEvery) : nullptr;
      hipEvent_t e0 = stamp ? c->tev[2 * (q / kTimeEvery)] : nullptr;
      hipEvent_t e1 = stamp ? c->tev[2 * (q / kTimeEvery) + 1] : nullptr;
      const double *pin = P[(p - 1) & 1];
      if (two && g.flow[0])
        hipExtLaunchKernelGGL(k_jacobi_pass2<false>, dim3(ts.B), dim3(ts.bx, ts.by), 0, c->stream,
                              e0, e1, 0, pg, jc, c->jb[0], c->jb[1], c->jb[2],
                              (const double *)c->b, pin, P[p & 1], c->js, (int32_t)p, ks);
      else if (two)
        hipExtLaunchKernelGGL(k_jacobi_pass2<true>, dim3(ts.B), dim3(ts.bx, ts.by), 0, c->stream,
                              e0, e1, 0, pg, jc, c->jb[0], c->jb[1], c->jb[2],
                              (const double *)c->b, pin, P[p & 1], c->js, (int32_t)p, ks);
      else
        hipExtLaunchKernelGGL(k_jacobi_pass, dim3(ts.B), dim3(ts.bx, ts.by), 0, c->stream,
                              e0, e1, 0, pg, jc, c->jb[0], c->jb[1], c->jb[2],
                              (const double *)c->b, pin, P[p & 1], c->js, (int32_t)p, ks);
    }
    PICARD_HIP(hipGetLastError());
    PICARD_HIP(hipMemcpyAsync(c->js_host, c->js, sizeof(JacobiState), hipMemcpyDeviceToHost,
                              c->stream));
    PICARD_HIP(hipStreamSynchronize(c->stream));
    if (c->timing) {
      for (int64_t q = 0; q < timed; ++q) {
        float ms = 0.0f;
        PICARD_HIP(hipEventElapsedTime(&ms, c->tev[2 * q], c->tev[2 * q + 1]));
        c->t_ms += ms;
      }
      c->t_launches += timed;
      c->t_sweeps += timed;
      PICARD_TRY(kstamp_collect(c, (int)timed));
    }
    fin = c->js_host->slot[(p - 1) & 1];
    if (fin.done) break;
    if (p > last + 1) {
      set_error("CalcEpotential: pass accounting out of step");
      return PICARD_ERR_STATE;
    }
    batch = 64;
  }
  c->last_sweeps = fin.result;
  k_jacobi_finish<<<nbn, kT, 0, c->stream>>>(g, c->jb[0], c->jb[1], c->jb[2], c->U, fin.result);
  PICARD_HIP(hipGetLastError());
  *sweeps = fin.result;
  // Running all 2*maxit+1 sweeps means every stop test failed; the reference
  // prints 'Maximum iteration reached!' on each sweep past maxit (:67-69).
  return fin.result == last ? PICARD_WARN_NOCONV : PICARD_OK;
}

int launch_calc_efield(picard_ctx *c) {
  k_efield<<<grid_blocks(c->gd.G, kT), kT, 0, c->stream>>>(c->gd, c->U, c->E);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int launch_bc_fields(picard_ctx *c) {
  k_bc_fields<<<grid_blocks(c->gd.G, kT), kT, 0, c->stream>>>(c->gd, c->E);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

}  // namespace picard
