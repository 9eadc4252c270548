/*
 * picard_oracle.c -- CPU restatement of Picard's PIC time step in FP64, in the
 * reference's evaluation order, for one MPI rank.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (no FMA contraction, so every
 * product and sum rounds exactly as written in the Fortran source).
 *
 * Each function cites the reference lines it restates.  On one rank every MPI
 * neighbour id equals myid, so each Isend/Irecv pair is a self-copy and each
 * Allreduce is the identity; the restatement performs those copies directly,
 * in the same z -> y -> x order and with the same "send buffer is a copy taken
 * before the receive is applied" semantics.
 *
 * Summation order: amdflang evaluates SUM(array-expression) as a sequential
 * column-major loop starting from 0 (checked with a probe of our own), so the
 * restatement sums sequentially in that order.
 */
#include "picard_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS0 8.854187817e-12                       /* BCpotentials.f90:19, CalcEpotential.f90:16 */
#define TAU 6.2831853071795864769252867665590      /* BCpotentials.f90:20 */

static inline int64_t gidx(const orc_grid *g, int i, int j, int k) {
  /* 1-based Fortran node (i,j,k) -> 0-based linear index */
  const int64_t X = g->nx + 2, Y = g->ny + 2;
  return (int64_t)(i - 1) + X * ((int64_t)(j - 1) + Y * (int64_t)(k - 1));
}

/* ------------------------------------------------------------------------ */
/* AdvancePosition.f90:15   r(1:3) = r(1:3) + r(4:6)*dt                       */
void orc_advance_position(double *rp, double dt, int64_t n) {
  for (int64_t p = 0; p < n; ++p) {
    double *r = rp + 8 * p;
    r[0] = r[0] + r[3] * dt;
    r[1] = r[1] + r[4] * dt;
    r[2] = r[2] + r[5] * dt;
  }
}

/* ------------------------------------------------------------------------ */
/* BCoutflow.f90:24-83                                                        */
int64_t orc_bc_outflow(const orc_grid *g, double *rp, int32_t *sp, int64_t *n_io,
                       int64_t *removed_slots) {
  int64_t n = *n_io, nrem = 0;
  const double xlo = 2.0 * g->lmin[0] - g->lmax[0], xhi = 2.0 * g->lmax[0] - g->lmin[0];
  const double ylo = 2.0 * g->lmin[1] - g->lmax[1], yhi = 2.0 * g->lmax[1] - g->lmin[1];
  const double zlo = 2.0 * g->lmin[2] - g->lmax[2], zhi = 2.0 * g->lmax[2] - g->lmin[2];
  for (int64_t pp = n; pp >= 1; --pp) {
    double *r = rp + 8 * (pp - 1);
    int del = 0;
    if (g->flow[0] && (r[0] < g->gmin[0] || r[0] > g->gmax[0])) del = 1;  /* :28-33 */
    if (g->flow[1] && (r[1] < g->gmin[1] || r[1] > g->gmax[1])) del = 1;  /* :37-42 */
    if (g->flow[2] && (r[2] < g->gmin[2] || r[2] > g->gmax[2])) del = 1;  /* :46-51 */
    if (r[0] < xlo || r[0] > xhi) del = 1;                                /* :56-58 */
    if (r[1] < ylo || r[1] > yhi) del = 1;                                /* :60-62 */
    if (r[2] < zlo || r[2] > zhi) del = 1;                                /* :64-66 */
    if (del) {                                                            /* :73-81 */
      if (removed_slots) removed_slots[nrem] = pp;
      ++nrem;
      if (pp != n) {
        memcpy(r, rp + 8 * (n - 1), 8 * sizeof(double));
        sp[pp - 1] = sp[n - 1];
      }
      n = n - 1;
    }
  }
  *n_io = n;
  return nrem;
}

/* ------------------------------------------------------------------------ */
/* ParticlesTransfer.f90:494-519 (idpz = idpy = idpx = myid)                  */
void orc_particles_transfer(const orc_grid *g, double *rp, int64_t n) {
  const double Lx = g->gmax[0] - g->gmin[0];
  const double Ly = g->gmax[1] - g->gmin[1];
  const double Lz = g->gmax[2] - g->gmin[2];
  for (int64_t pp = n; pp >= 1; --pp) {
    double *r = rp + 8 * (pp - 1);
    if (r[2] < g->gmin[2]) r[2] = r[2] + Lz;
    if (r[2] > g->gmax[2]) r[2] = r[2] - Lz;
    if (r[1] < g->gmin[1]) r[1] = r[1] + Ly;
    if (r[1] > g->gmax[1]) r[1] = r[1] - Ly;
    if (r[0] < g->gmin[0]) r[0] = r[0] + Lx;
    if (r[0] > g->gmax[0]) r[0] = r[0] - Lx;
  }
}

/* ------------------------------------------------------------------------ */
/* CIC location, ParticlesGridify.f90:70-77 and AdvanceVelocity.f90:44-51:
 * rest = dmod(x - Lmin + dx/2, dx) and idx = 1 + idint((x - Lmin)/dx + 1/2)
 * are evaluated independently (flang's dmod equals C fmod, which is exact). */
static inline void cic_locate(const orc_grid *g, const double *r, double rest[3],
                              double drest[3], int idx[3]) {
  for (int c = 0; c < 3; ++c) {
    rest[c] = fmod(r[c] - g->lmin[c] + 0.5 * g->dxyz[c], g->dxyz[c]);
    drest[c] = g->dxyz[c] - rest[c];
    idx[c] = 1 + (int)((r[c] - g->lmin[c]) / g->dxyz[c] + 0.5);
  }
}

static inline int cic_in_bounds(const orc_grid *g, const int idx[3]) {
  return idx[0] >= 1 && idx[1] >= 1 && idx[2] >= 1 && idx[0] <= g->nx + 1 &&
         idx[1] <= g->ny + 1 && idx[2] <= g->nz + 1;
}

/* Guard fold for one dimension, ParticlesGridify.f90:129-285 on one rank.
 * dim 0 = x, 1 = y, 2 = z.  Works on whole planes of N(S, ...). */
static void gridify_fold_n(const orc_grid *g, double *N, int dim, int S) {
  const int X = g->nx + 2, Y = g->ny + 2, Z = g->nz + 2;
  const int n = dim == 0 ? g->nx : (dim == 1 ? g->ny : g->nz);
  const int e1 = dim == 0 ? Y : X;          /* extents of the plane */
  const int e2 = dim == 2 ? Y : Z;
  for (int b = 1; b <= e2; ++b) {
    for (int a = 1; a <= e1; ++a) {
      for (int s = 0; s < S; ++s) {
        int64_t at[4];
        /* plane positions 1, 2, n+1, n+2 along dim */
        const int pos[4] = {1, 2, n + 1, n + 2};
        for (int q = 0; q < 4; ++q) {
          int i, j, k;
          if (dim == 0) { i = pos[q]; j = a; k = b; }
          else if (dim == 1) { i = a; j = pos[q]; k = b; }
          else { i = a; j = b; k = pos[q]; }
          at[q] = (int64_t)S * gidx(g, i, j, k) + s;
        }
        const double o1 = N[at[0]], o2 = N[at[1]], on1 = N[at[2]], on2 = N[at[3]];
        if (g->flow[dim]) {
          /* :149-163  fold guard into first interior, zero the guard */
          N[at[1]] = o2 + o1;
          N[at[0]] = 0.0;
          N[at[2]] = on1 + on2;
          N[at[3]] = 0.0;
        } else {
          /* :166-169  periodic: add the self-exchanged 2-plane slab */
          N[at[0]] = o1 + on1;
          N[at[1]] = o2 + on2;
          N[at[2]] = on1 + o1;
          N[at[3]] = on2 + o2;
        }
      }
    }
  }
}

static void gridify_fold(const orc_grid *g, double *N, int dim) {
  gridify_fold_n(g, N, dim, g->species);
}

/* ParticlesGridify.f90:56-293 */
int64_t orc_particles_gridify(const orc_grid *g, double *N, double *C, const double *rp,
                              const int32_t *sp, int64_t n, const double charge[8],
                              const double n2p[8]) {
  const int S = g->species;
  const int64_t G = (int64_t)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
  const double dV = g->dxyz[0] * g->dxyz[1] * g->dxyz[2];
  int64_t skipped = 0;
  for (int64_t q = 0; q < (int64_t)S * G; ++q) N[q] = 0.0;   /* :59 */
  for (int64_t q = 0; q < G; ++q) C[q] = 0.0;                /* :60 */
  for (int64_t pp = n; pp >= 1; --pp) {                      /* :65 */
    const double *r = rp + 8 * (pp - 1);
    double rest[3], drest[3];
    int idx[3];
    cic_locate(g, r, rest, drest, idx);
    if (!cic_in_bounds(g, idx)) { ++skipped; continue; }
    const double F1 = drest[0] * drest[1] * drest[2] / dV;   /* :81-88 */
    const double F2 = rest[0] * drest[1] * drest[2] / dV;
    const double F3 = drest[0] * rest[1] * drest[2] / dV;
    const double F4 = drest[0] * drest[1] * rest[2] / dV;
    const double F5 = rest[0] * rest[1] * drest[2] / dV;
    const double F6 = drest[0] * rest[1] * rest[2] / dV;
    const double F7 = rest[0] * drest[1] * rest[2] / dV;
    const double F8 = rest[0] * rest[1] * rest[2] / dV;
    const int s = sp[pp - 1];
    const double w = n2p[s - 1];
    const int i = idx[0], j = idx[1], k = idx[2];
    double *Ns = N + (s - 1);
#define NAT(a, b, c) Ns[(int64_t)S * gidx(g, a, b, c)]
    NAT(i + 0, j + 0, k + 0) = NAT(i + 0, j + 0, k + 0) + F1 * w;   /* :98-121 */
    NAT(i + 1, j + 0, k + 0) = NAT(i + 1, j + 0, k + 0) + F2 * w;
    NAT(i + 0, j + 1, k + 0) = NAT(i + 0, j + 1, k + 0) + F3 * w;
    NAT(i + 0, j + 0, k + 1) = NAT(i + 0, j + 0, k + 1) + F4 * w;
    NAT(i + 1, j + 1, k + 0) = NAT(i + 1, j + 1, k + 0) + F5 * w;
    NAT(i + 0, j + 1, k + 1) = NAT(i + 0, j + 1, k + 1) + F6 * w;
    NAT(i + 1, j + 0, k + 1) = NAT(i + 1, j + 0, k + 1) + F7 * w;
    NAT(i + 1, j + 1, k + 1) = NAT(i + 1, j + 1, k + 1) + F8 * w;
#undef NAT
  }
  gridify_fold(g, N, 2);   /* Z BOUNDARY :127-178 */
  gridify_fold(g, N, 1);   /* Y BOUNDARY :181-232 */
  gridify_fold(g, N, 0);   /* X BOUNDARY :234-285 */
  for (int s = 0; s < S; ++s)                                /* :288-293 */
    for (int64_t q = 0; q < G; ++q) C[q] = C[q] + N[(int64_t)S * q + s] * charge[s];
  return skipped;
}

/* ------------------------------------------------------------------------ */
/* CalcDistance.f90:25-91                                                     */
static inline double inv_len(double a, double b, double c) {
  return 1.0 / sqrt(a * a + b * b + c * c);   /* dsqrt( x**2 + y**2 + z**2 )**(-1) */
}

void orc_calc_distance(double *D, int32_t Nx, int32_t Ny, int32_t Nz, const double dxyz[3],
                       const int32_t flow[3]) {
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1, DZ = 2 * (int64_t)Nz + 1;
  const int xf = flow[0] != 0, yf = flow[1] != 0, zf = flow[2] != 0;
  const double hx = dxyz[0], hy = dxyz[1], hz = dxyz[2];
  memset(D, 0, sizeof(double) * DX * DY * DZ);
#define DAT(a, b, c) D[(int64_t)((a) - 1) + DX * ((int64_t)((b) - 1) + DY * (int64_t)((c) - 1))]
  for (int kk = 1; kk <= Nz + 1; ++kk)
    for (int jj = 1; jj <= Ny + 1; ++jj)
      for (int ii = 1; ii <= Nx + 1; ++ii) {
        const int pi = Nx + 1 + (ii - 1), pj = Ny + 1 + (jj - 1), pk = Nz + 1 + (kk - 1);
        const int ni = Nx + 1 - (ii - 1), nj = Ny + 1 - (jj - 1), nk = Nz + 1 - (kk - 1);
        const double di = (double)(ii - 1), dj = (double)(jj - 1), dk = (double)(kk - 1);
        double t;
        if (ii == 1 && jj == 1 && kk == 1) t = 0.0;                 /* :42-46 */
        else t = inv_len(di * hx, dj * hy, dk * hz);
        const double mx = ((double)Nx - di) * hx, my = ((double)Ny - dj) * hy,
                     mz = ((double)Nz - dk) * hz;
        if (xf) {                                                   /* :48-57 */
          if (!yf && !zf) {
            t = t + inv_len(di * hx, my, dk * hz);
            t = t + inv_len(di * hx, dj * hy, mz);
            t = t + inv_len(di * hx, my, mz);
          } else if (!yf) {
            t = t + inv_len(di * hx, my, dk * hz);
          } else if (!zf) {
            t = t + inv_len(di * hx, dj * hy, mz);
          }
        } else if (yf) {                                            /* :58-65 */
          if (!zf) {
            t = t + inv_len(mx, dj * hy, dk * hz);
            t = t + inv_len(di * hx, dj * hy, mz);
            t = t + inv_len(mx, dj * hy, mz);
          } else {
            t = t + inv_len(mx, dj * hy, dk * hz);
          }
        } else if (zf) {                                            /* :66-69 */
          t = t + inv_len(mx, dj * hy, dk * hz);
          t = t + inv_len(di * hx, my, dk * hz);
          t = t + inv_len(mx, my, dk * hz);
        } else {                                                    /* :70-77 */
          t = t + inv_len(mx, dj * hy, dk * hz);
          t = t + inv_len(di * hx, my, dk * hz);
          t = t + inv_len(di * hx, dj * hy, mz);
          t = t + inv_len(mx, my, dk * hz);
          t = t + inv_len(di * hx, my, mz);
          t = t + inv_len(mx, dj * hy, mz);
          t = t + inv_len(mx, my, mz);
        }
        const double v = 1.0 * t;                                   /* :80-87 */
        DAT(pi, pj, pk) = v; DAT(ni, pj, pk) = v; DAT(pi, nj, pk) = v; DAT(pi, pj, nk) = v;
        DAT(ni, nj, pk) = v; DAT(pi, nj, nk) = v; DAT(ni, pj, nk) = v; DAT(ni, nj, nk) = v;
      }
#undef DAT
}

/* ------------------------------------------------------------------------ */
/* BCpotentials.f90:36-141, one rank: bi = bj = bk = 0, N_local = N.          */
void orc_bc_potentials(const orc_grid *g, double *U, const double *C, const double *D) {
  const int Nx = g->nx, Ny = g->ny, Nz = g->nz;
  const int64_t G = (int64_t)(Nx + 2) * (Ny + 2) * (Nz + 2);
  const int64_t DX = 2 * (int64_t)Nx + 1, DY = 2 * (int64_t)Ny + 1;
  const double kfac = 1.0 / (2.0 * TAU * EPS0);          /* (2.0d0*tau*eps0)**(-1) */
#define DAT(a, b, c) D[(int64_t)((a) - 1) + DX * ((int64_t)((b) - 1) + DY * (int64_t)((c) - 1))]
  for (int64_t q = 0; q < G; ++q) U[q] = 0.0;            /* :36 */
  /* face sum: sum( C(2:Nx+1,2:Ny+1,2:Nz+1) * D(di:,dj:,dk:) ), column-major */
#define FACE_SUM(out, di, dj, dk)                                        \
  do {                                                                   \
    double s_ = 0.0;                                                     \
    for (int k_ = 1; k_ <= Nz; ++k_)                                     \
      for (int j_ = 1; j_ <= Ny; ++j_)                                   \
        for (int i_ = 1; i_ <= Nx; ++i_)                                 \
          s_ = s_ + C[gidx(g, i_ + 1, j_ + 1, k_ + 1)] *                 \
                        DAT((di) + i_ - 1, (dj) + j_ - 1, (dk) + k_ - 1);\
    (out) = s_;                                                          \
  } while (0)
  if (g->flow[0]) {                                      /* :51-79 */
    for (int kk = 1; kk <= Nz; ++kk)
      for (int jj = 1; jj <= Ny; ++jj) {
        const int dj = Ny + 1 + (1 - jj), dk = Nz + 1 + (1 - kk);
        double u0, u1;
        FACE_SUM(u0, Nx + 1 + 1, dj, dk);
        FACE_SUM(u1, Nx + 1 - Nx, dj, dk);
        U[gidx(g, 1, jj + 1, kk + 1)] = u0 * kfac;
        U[gidx(g, Nx + 2, jj + 1, kk + 1)] = u1 * kfac;
      }
  }
  if (g->flow[1]) {                                      /* :82-110 */
    for (int kk = 1; kk <= Nz; ++kk)
      for (int ii = 1; ii <= Nx; ++ii) {
        const int di = Nx + 1 + (1 - ii), dk = Nz + 1 + (1 - kk);
        double u0, u1;
        FACE_SUM(u0, di, Ny + 1 + 1, dk);
        FACE_SUM(u1, di, Ny + 1 - Ny, dk);
        U[gidx(g, ii + 1, 1, kk + 1)] = u0 * kfac;
        U[gidx(g, ii + 1, Ny + 2, kk + 1)] = u1 * kfac;
      }
  }
  if (g->flow[2]) {                                      /* :113-141 */
    for (int jj = 1; jj <= Ny; ++jj)
      for (int ii = 1; ii <= Nx; ++ii) {
        const int di = Nx + 1 + (1 - ii), dj = Ny + 1 + (1 - jj);
        double u0, u1;
        FACE_SUM(u0, di, dj, Nz + 1 + 1);
        FACE_SUM(u1, di, dj, Nz + 1 - Nz);
        U[gidx(g, ii + 1, jj + 1, 1)] = u0 * kfac;
        U[gidx(g, ii + 1, jj + 1, Nz + 2)] = u1 * kfac;
      }
  }
#undef FACE_SUM
#undef DAT
}

/* ------------------------------------------------------------------------ */
/* Scalar halo exchange of U for one dimension, CalcEpotential.f90:100-239 on
 * one rank: flow faces keep their values, periodic faces copy whole planes. */
static void halo_scalar(const orc_grid *g, double *U, int dim) {
  if (g->flow[dim]) return;
  const int X = g->nx + 2, Y = g->ny + 2, Z = g->nz + 2;
  const int n = dim == 0 ? g->nx : (dim == 1 ? g->ny : g->nz);
  const int e1 = dim == 0 ? Y : X, e2 = dim == 2 ? Y : Z;
  for (int b = 1; b <= e2; ++b)
    for (int a = 1; a <= e1; ++a) {
      int64_t lo, lo_src, hi, hi_src;
      if (dim == 0) {
        lo = gidx(g, 1, a, b); lo_src = gidx(g, n + 1, a, b);
        hi = gidx(g, n + 2, a, b); hi_src = gidx(g, 2, a, b);
      } else if (dim == 1) {
        lo = gidx(g, a, 1, b); lo_src = gidx(g, a, n + 1, b);
        hi = gidx(g, a, n + 2, b); hi_src = gidx(g, a, 2, b);
      } else {
        lo = gidx(g, a, b, 1); lo_src = gidx(g, a, b, n + 1);
        hi = gidx(g, a, b, n + 2); hi_src = gidx(g, a, b, 2);
      }
      /* send buffers are copies taken before either receive is applied */
      const double vneg = U[hi_src], vpos = U[lo_src];
      U[lo] = vpos;
      U[hi] = vneg;
    }
}

static double pairwise_sum(const double *v, int64_t n) {
  if (n <= 8) {
    double s = 0.0;
    for (int64_t q = 0; q < n; ++q) s = s + v[q];
    return s;
  }
  const int64_t h = n / 2;
  return pairwise_sum(v, h) + pairwise_sum(v + h, n - h);
}

/* CalcEpotential.f90:28-242 */
static int64_t calc_epotential_impl(const orc_grid *g, double *U, const double *C,
                                    double maxerr, int32_t maxit, double *final_ratio,
                                    int pairwise) {
  const int Nx = g->nx, Ny = g->ny, Nz = g->nz;
  const int64_t X = Nx + 2, Y = Ny + 2;
  const int64_t G = X * Y * (Nz + 2);
  const double dx = g->dxyz[0], dy = g->dxyz[1], dz = g->dxyz[2];
  const double dV = dx * dy * dz;
  double *b = (double *)calloc((size_t)G, sizeof(double));
  double *res = (double *)calloc((size_t)G, sizeof(double));
  double *unew = (double *)malloc(sizeof(double) * (size_t)G);
  double *sq = pairwise ? (double *)malloc(sizeof(double) * (size_t)Nx * Ny * Nz) : NULL;
  /* :44-47 */
  const double den = 1.0 / ((dx * dy) * (dx * dy) + (dx * dz) * (dx * dz) + (dy * dz) * (dy * dz));
  const double ax = 0.5 * den * ((dy * dz) * (dy * dz));
  const double ay = 0.5 * den * ((dx * dz) * (dx * dz));
  const double az = 0.5 * den * ((dx * dy) * (dx * dy));
  const double af = 0.5 * den * ((dx * dy * dz) * (dx * dy * dz));
  const double binv = 1.0 / (dV * EPS0);
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  for (int64_t q = 0; q < G; ++q) b[q] = -C[q] * binv;      /* :50 */
  double b_global = 0.0;                                     /* :52 */
  {
    int64_t m = 0;
    for (int k = 2; k <= Nz + 1; ++k)
      for (int j = 2; j <= Ny + 1; ++j)
        for (int i = 2; i <= Nx + 1; ++i) {
          const double v = b[gidx(g, i, j, k)];
          if (pairwise) sq[m++] = v * v;
          else b_global = b_global + v * v;
        }
    if (pairwise) b_global = pairwise_sum(sq, m);
  }
  const double b_sqrt = sqrt(b_global);                      /* :54 */
  double r_global = 1.0 * b_global;                          /* :55 */
  int64_t sweeps = 0;
  const int64_t last = 2 * (int64_t)maxit + 1;
  for (int64_t ss = 1; ss <= last; ++ss) {                   /* :58 */
    if (ss <= maxit) {
      if (sqrt(r_global) <= maxerr * b_sqrt) break;          /* :60-62 */
    }
    /* :74-80  pure Jacobi: the right side is evaluated in full first */
    for (int k = 2; k <= Nz + 1; ++k)
      for (int j = 2; j <= Ny + 1; ++j)
        for (int i = 2; i <= Nx + 1; ++i) {
          const int64_t c = gidx(g, i, j, k);
          unew[c] = ax * (U[c + 1] + U[c - 1]) + ay * (U[c + X] + U[c - X]) +
                    az * (U[c + X * Y] + U[c - X * Y]) - af * b[c];
        }
    for (int k = 2; k <= Nz + 1; ++k)
      for (int j = 2; j <= Ny + 1; ++j)
        for (int i = 2; i <= Nx + 1; ++i) {
          const int64_t c = gidx(g, i, j, k);
          U[c] = unew[c];
        }
    /* :83-92  residual of the new iterate with the halos of the old one */
    double r_local = 0.0;
    int64_t m = 0;
    for (int k = 2; k <= Nz + 1; ++k)
      for (int j = 2; j <= Ny + 1; ++j)
        for (int i = 2; i <= Nx + 1; ++i) {
          const int64_t c = gidx(g, i, j, k);
          const double u2 = 2.0 * U[c];
          res[c] = b[c] - ((U[c + 1] + U[c - 1] - u2) / dx2 + (U[c + X] + U[c - X] - u2) / dy2 +
                           (U[c + X * Y] + U[c - X * Y] - u2) / dz2);
          if (pairwise) sq[m++] = res[c] * res[c];
          else r_local = r_local + res[c] * res[c];          /* :95 */
        }
    if (pairwise) r_local = pairwise_sum(sq, m);
    r_global = r_local;                                      /* :96 */
    halo_scalar(g, U, 2);                                    /* :100-145 */
    halo_scalar(g, U, 1);                                    /* :147-192 */
    halo_scalar(g, U, 0);                                    /* :194-239 */
    sweeps = ss;
  }
  if (final_ratio) *final_ratio = b_sqrt > 0.0 ? sqrt(r_global) / b_sqrt : 0.0;
  free(b); free(res); free(unew); free(sq);
  return sweeps;
}

int64_t orc_calc_epotential(const orc_grid *g, double *U, const double *C, double maxerr,
                            int32_t maxit, double *final_ratio) {
  return calc_epotential_impl(g, U, C, maxerr, maxit, final_ratio, 0);
}

int64_t orc_calc_epotential_pairwise(const orc_grid *g, double *U, const double *C,
                                     double maxerr, int32_t maxit) {
  return calc_epotential_impl(g, U, C, maxerr, maxit, NULL, 1);
}

/* ------------------------------------------------------------------------ */
/* Vector (3-component) halo exchange, CalcEfield.f90:92-228 on one rank.     */
static void halo_vec3(const orc_grid *g, double *E, int dim) {
  if (g->flow[dim]) return;
  const int X = g->nx + 2, Y = g->ny + 2, Z = g->nz + 2;
  const int n = dim == 0 ? g->nx : (dim == 1 ? g->ny : g->nz);
  const int e1 = dim == 0 ? Y : X, e2 = dim == 2 ? Y : Z;
  for (int b = 1; b <= e2; ++b)
    for (int a = 1; a <= e1; ++a) {
      int64_t lo, lo_src, hi, hi_src;
      if (dim == 0) {
        lo = gidx(g, 1, a, b); lo_src = gidx(g, n + 1, a, b);
        hi = gidx(g, n + 2, a, b); hi_src = gidx(g, 2, a, b);
      } else if (dim == 1) {
        lo = gidx(g, a, 1, b); lo_src = gidx(g, a, n + 1, b);
        hi = gidx(g, a, n + 2, b); hi_src = gidx(g, a, 2, b);
      } else {
        lo = gidx(g, a, b, 1); lo_src = gidx(g, a, b, n + 1);
        hi = gidx(g, a, b, n + 2); hi_src = gidx(g, a, b, 2);
      }
      for (int c = 0; c < 3; ++c) {
        const double vneg = E[3 * hi_src + c], vpos = E[3 * lo_src + c];
        E[3 * lo + c] = vpos;
        E[3 * hi + c] = vneg;
      }
    }
}

/* CalcEfield.f90:27-228 */
void orc_calc_efield(const orc_grid *g, double *E, const double *U) {
  const int Nx = g->nx, Ny = g->ny, Nz = g->nz;
  const int64_t X = Nx + 2, Y = Ny + 2, G = X * Y * (Nz + 2);
  for (int64_t q = 0; q < 3 * G; ++q) E[q] = 0.0;                   /* :27 */
  for (int k = 2; k <= Nz + 1; ++k)
    for (int j = 2; j <= Ny + 1; ++j)
      for (int i = 2; i <= Nx + 1; ++i) {
        const int64_t c = gidx(g, i, j, k);
        E[3 * c + 2] = -(U[c + X * Y] - U[c - X * Y]) * 0.5 / g->dxyz[2];  /* :29-30 */
        E[3 * c + 1] = -(U[c + X] - U[c - X]) * 0.5 / g->dxyz[1];          /* :31-32 */
        E[3 * c + 0] = -(U[c + 1] - U[c - 1]) * 0.5 / g->dxyz[0];          /* :33-34 */
      }
  halo_vec3(g, E, 2);
  halo_vec3(g, E, 1);
  halo_vec3(g, E, 0);
}

/* BCfields.f90:31-90, one rank: every owned flow face gets its guard zeroed. */
void orc_bc_fields(const orc_grid *g, double *E) {
  const int X = g->nx + 2, Y = g->ny + 2, Z = g->nz + 2;
  for (int k = 1; k <= Z; ++k)
    for (int j = 1; j <= Y; ++j)
      for (int i = 1; i <= X; ++i) {
        int zero = 0;
        if (g->flow[2] && (k == Z || k == 1)) zero = 1;
        if (g->flow[1] && (j == Y || j == 1)) zero = 1;
        if (g->flow[0] && (i == X || i == 1)) zero = 1;
        if (zero) {
          const int64_t c = gidx(g, i, j, k);
          E[3 * c] = 0.0; E[3 * c + 1] = 0.0; E[3 * c + 2] = 0.0;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* AdvanceVelocity.f90:34-121                                                 */
void orc_advance_velocity(const orc_grid *g, double *rp, const int32_t *sp, const double *E,
                          const double B0[4], double dt, const double ve[4],
                          const double eta[8], const double tB[32], const double sB[32],
                          int64_t n, const int32_t fluid[8], const int32_t gyration[8]) {
  const double dV = g->dxyz[0] * g->dxyz[1] * g->dxyz[2];           /* :34 */
  double etadt2[8];
  for (int s = 0; s < 8; ++s) etadt2[s] = 0.5 * dt * eta[s];        /* :35 */
  for (int64_t pp = n; pp >= 1; --pp) {                             /* :39 */
    double *r = rp + 8 * (pp - 1);
    double rest[3], drest[3];
    int idx[3];
    cic_locate(g, r, rest, drest, idx);
    if (!cic_in_bounds(g, idx)) continue;
    const double F1 = drest[0] * drest[1] * drest[2];                /* :65-72 */
    const double F2 = rest[0] * drest[1] * drest[2];
    const double F3 = drest[0] * rest[1] * drest[2];
    const double F4 = drest[0] * drest[1] * rest[2];
    const double F5 = rest[0] * rest[1] * drest[2];
    const double F6 = drest[0] * rest[1] * rest[2];
    const double F7 = rest[0] * drest[1] * rest[2];
    const double F8 = rest[0] * rest[1] * rest[2];
    const int i = idx[0], j = idx[1], k = idx[2];
    double Ep[3];
    for (int c = 0; c < 3; ++c) {                                   /* :74-78 */
      const double e111 = E[3 * gidx(g, i, j, k) + c], e211 = E[3 * gidx(g, i + 1, j, k) + c];
      const double e121 = E[3 * gidx(g, i, j + 1, k) + c], e112 = E[3 * gidx(g, i, j, k + 1) + c];
      const double e221 = E[3 * gidx(g, i + 1, j + 1, k) + c];
      const double e122 = E[3 * gidx(g, i, j + 1, k + 1) + c];
      const double e212 = E[3 * gidx(g, i + 1, j, k + 1) + c];
      const double e222 = E[3 * gidx(g, i + 1, j + 1, k + 1) + c];
      Ep[c] = e111 * F1 + e211 * F2 + e121 * F3 + e112 * F4 + e221 * F5 + e122 * F6 +
              e212 * F7 + e222 * F8;
      Ep[c] = Ep[c] / dV;
    }
    const int s = sp[pp - 1] - 1;
    if (fluid[s]) {                                                 /* :85-96 */
      if (gyration[s]) {
        const double B1 = B0[0], B2 = B0[1], B3 = B0[2];
        const double vdb = r[3] * B1 + r[4] * B2 + r[5] * B3;
        const double vm1 = vdb * B1, vm2 = vdb * B2, vm3 = vdb * B3;
        const double vd1 = vm1 + (Ep[1] * B3 - Ep[2] * B2);
        const double vd2 = vm2 + (Ep[2] * B1 - Ep[0] * B3);
        const double vd3 = vm3 + (Ep[0] * B2 - Ep[1] * B1);
        const double inv = 1.0 / (B1 * B1 + B2 * B2 + B3 * B3);
        r[3] = vd1 * inv; r[4] = vd2 * inv; r[5] = vd3 * inv;
      }
    } else {
      const double h = etadt2[s];
      const double *t = tB + 4 * s, *q = sB + 4 * s;
      const double vm1 = r[3] - ve[0] + h * Ep[0];                  /* :102 */
      const double vm2 = r[4] - ve[1] + h * Ep[1];
      const double vm3 = r[5] - ve[2] + h * Ep[2];
      const double vd1 = vm1 + vm2 * t[2] - vm3 * t[1];             /* :105-107 */
      const double vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
      const double vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
      const double vp1 = vm1 + vd2 * q[2] - vd3 * q[1];             /* :112-114 */
      const double vp2 = vm2 + vd3 * q[0] - vd1 * q[2];
      const double vp3 = vm3 + vd1 * q[1] - vd2 * q[0];
      r[3] = vp1 + ve[0] + h * Ep[0];                               /* :117 */
      r[4] = vp2 + ve[1] + h * Ep[1];
      r[5] = vp3 + ve[2] + h * Ep[2];
    }
  }
}

/* ---- diagnostics (picard.f90:649-672, one rank) -------------------------
 * out[0] = uP = 0.5 * SUM_pp n2p(s) mass(s) (v1^2 + v2^2 + v3^2)   (:658-661)
 * out[1] = uE = SUM_interior |E + E0|^2 * 0.5 * eps0 * dV          (:655-657)
 * out[2] = uB = SUM_interior,c B(c)^2 * 0.5 / mu0 * dV, B = B0     (:653, :420-422)
 * Each SUM runs sequentially in Fortran array order (first index fastest). */
void orc_diagnostics(const orc_grid *g, const double *rp, const int32_t *sp, int64_t n,
                     const double *E, const double *n2p, const double *mass, const double *E0,
                     const double *B0, double out[3]) {
  const double eps0 = 8.854187817e-12, mu0 = 1.25663706143591729538505735331180e-6;
  const double dV = g->dxyz[0] * g->dxyz[1] * g->dxyz[2];
  double up = 0.0;
  for (int64_t pp = 0; pp < n; ++pp) {
    const double *r = rp + 8 * pp;
    const int s = sp[pp] - 1;
    up = up + n2p[s] * mass[s] * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
  }
  out[0] = up * 0.5;
  const int X = g->nx + 2, Y = g->ny + 2;
  double ue = 0.0, ub = 0.0;
  for (int k = 2; k <= g->nz + 1; ++k)
    for (int j = 2; j <= g->ny + 1; ++j)
      for (int i = 2; i <= g->nx + 1; ++i) {
        const double *e = E + 3 * ((int64_t)(i - 1) + (int64_t)X * ((j - 1) + (int64_t)Y * (k - 1)));
        const double a = e[0] + E0[0], b = e[1] + E0[1], c = e[2] + E0[2];
        ue = ue + (a * a + b * b + c * c);
        for (int q = 0; q < 3; ++q) ub = ub + B0[q] * B0[q];
      }
  out[1] = ue * 0.5 * eps0 * dV;
  out[2] = ub * 0.5 / mu0 * dV;
}

/* ---- AdvanceVelocityImp (mode 1, AdvanceVelocity.f90:240-347) and
 * AdvanceVelocityExp (mode 2, :128-238): same gather and fluid branch as
 * AdvanceVelocity; the kick and a norm-preserving rotation with etadt2 and
 * tB scaled by 2 deltat/dt (:156-157, :267-268). */
void orc_advance_velocity_ie(int mode, const orc_grid *g, double *rp, const int32_t *sp,
                             const double *E, const double B0[4], double deltat, double dt,
                             const double ve[4], const double eta[8], const double tB[32],
                             int64_t n, const int32_t fluid[8], const int32_t gyration[8]) {
  const double dV = g->dxyz[0] * g->dxyz[1] * g->dxyz[2];
  double etadt2[8], tBt[32];
  for (int s = 0; s < 8; ++s) etadt2[s] = 0.5 * dt * eta[s] * 2.0 * deltat / dt;
  for (int q = 0; q < 32; ++q) tBt[q] = tB[q] * 2.0 * deltat / dt;
  for (int64_t pp = n; pp >= 1; --pp) {
    double *r = rp + 8 * (pp - 1);
    double rest[3], drest[3];
    int idx[3];
    cic_locate(g, r, rest, drest, idx);
    if (!cic_in_bounds(g, idx)) continue;
    const double F1 = drest[0] * drest[1] * drest[2];
    const double F2 = rest[0] * drest[1] * drest[2];
    const double F3 = drest[0] * rest[1] * drest[2];
    const double F4 = drest[0] * drest[1] * rest[2];
    const double F5 = rest[0] * rest[1] * drest[2];
    const double F6 = drest[0] * rest[1] * rest[2];
    const double F7 = rest[0] * drest[1] * rest[2];
    const double F8 = rest[0] * rest[1] * rest[2];
    const int i = idx[0], j = idx[1], k = idx[2];
    double Ep[3];
    for (int c = 0; c < 3; ++c) {
      const double e111 = E[3 * gidx(g, i, j, k) + c], e211 = E[3 * gidx(g, i + 1, j, k) + c];
      const double e121 = E[3 * gidx(g, i, j + 1, k) + c], e112 = E[3 * gidx(g, i, j, k + 1) + c];
      const double e221 = E[3 * gidx(g, i + 1, j + 1, k) + c];
      const double e122 = E[3 * gidx(g, i, j + 1, k + 1) + c];
      const double e212 = E[3 * gidx(g, i + 1, j, k + 1) + c];
      const double e222 = E[3 * gidx(g, i + 1, j + 1, k + 1) + c];
      Ep[c] = e111 * F1 + e211 * F2 + e121 * F3 + e112 * F4 + e221 * F5 + e122 * F6 +
              e212 * F7 + e222 * F8;
      Ep[c] = Ep[c] / dV;
    }
    const int s = sp[pp - 1] - 1;
    if (fluid[s]) {
      if (gyration[s]) {
        const double B1 = B0[0], B2 = B0[1], B3 = B0[2];
        const double vdb = r[3] * B1 + r[4] * B2 + r[5] * B3;
        const double vm1 = vdb * B1, vm2 = vdb * B2, vm3 = vdb * B3;
        const double vd1 = vm1 + (Ep[1] * B3 - Ep[2] * B2);
        const double vd2 = vm2 + (Ep[2] * B1 - Ep[0] * B3);
        const double vd3 = vm3 + (Ep[0] * B2 - Ep[1] * B1);
        const double inv = 1.0 / (B1 * B1 + B2 * B2 + B3 * B3);
        r[3] = vd1 * inv; r[4] = vd2 * inv; r[5] = vd3 * inv;
      }
      continue;
    }
    const double h = etadt2[s];
    const double *t = tBt + 4 * s;
    double vm1 = r[3] - ve[0], vm2 = r[4] - ve[1], vm3 = r[5] - ve[2];
    if (mode == 1) {                                                /* :297 */
      vm1 = vm1 + h * Ep[0];
      vm2 = vm2 + h * Ep[1];
      vm3 = vm3 + h * Ep[2];
    }
    double vd1 = vm1 + vm2 * t[2] - vm3 * t[1];                     /* :300-302 */
    double vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
    double vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
    const double sd = vd1 * vd1 + vd2 * vd2 + vd3 * vd3;
    if (sd > 0.0) {                                                 /* :305-307 */
      const double f = sqrt((vm1 * vm1 + vm2 * vm2 + vm3 * vm3) / sd);
      vd1 = vd1 * f; vd2 = vd2 * f; vd3 = vd3 * f;
    }
    if (mode == 1) {                                                /* :310 */
      r[3] = vd1 + ve[0]; r[4] = vd2 + ve[1]; r[5] = vd3 + ve[2];
    } else {                                                        /* :222 */
      r[3] = vd1 + ve[0] + h * Ep[0];
      r[4] = vd2 + ve[1] + h * Ep[1];
      r[5] = vd3 + ve[2] + h * Ep[2];
    }
  }
}

/* ---- ParticlesGridifyAll.f90:56-452 (one rank): ParticlesGridify's N and C
 * plus F(s, c, node) = F(s, c, node) + Fk * n2p(s) * v_c in the same loop
 * (:126-212), then the same z, y, x guard folds on N and F (:238-420). */
int64_t orc_particles_gridify_all(const orc_grid *g, double *F, double *N, double *C,
                                  const double *rp, const int32_t *sp, int64_t n,
                                  const double charge[8], const double n2p[8]) {
  const int S = g->species;
  const int64_t G = (int64_t)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
  const double dV = g->dxyz[0] * g->dxyz[1] * g->dxyz[2];
  int64_t skipped = 0;
  for (int64_t q = 0; q < 3 * (int64_t)S * G; ++q) F[q] = 0.0;   /* :61 */
  for (int64_t q = 0; q < (int64_t)S * G; ++q) N[q] = 0.0;
  for (int64_t q = 0; q < G; ++q) C[q] = 0.0;
  for (int64_t pp = n; pp >= 1; --pp) {
    const double *r = rp + 8 * (pp - 1);
    double rest[3], drest[3];
    int idx[3];
    cic_locate(g, r, rest, drest, idx);
    if (!cic_in_bounds(g, idx)) { ++skipped; continue; }
    double Fk[8];
    Fk[0] = drest[0] * drest[1] * drest[2] / dV;
    Fk[1] = rest[0] * drest[1] * drest[2] / dV;
    Fk[2] = drest[0] * rest[1] * drest[2] / dV;
    Fk[3] = drest[0] * drest[1] * rest[2] / dV;
    Fk[4] = rest[0] * rest[1] * drest[2] / dV;
    Fk[5] = drest[0] * rest[1] * rest[2] / dV;
    Fk[6] = rest[0] * drest[1] * rest[2] / dV;
    Fk[7] = rest[0] * rest[1] * rest[2] / dV;
    /* node offsets of F1..F8 */
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    const int s = sp[pp - 1] - 1;
    const double w = n2p[s];
    for (int q = 0; q < 8; ++q) {
      const int64_t node = gidx(g, idx[0] + off[q][0], idx[1] + off[q][1], idx[2] + off[q][2]);
      N[(int64_t)S * node + s] = N[(int64_t)S * node + s] + Fk[q] * w;
    }
    for (int c = 0; c < 3; ++c)
      for (int q = 0; q < 8; ++q) {
        const int64_t node = gidx(g, idx[0] + off[q][0], idx[1] + off[q][1], idx[2] + off[q][2]);
        double *f = F + (int64_t)3 * S * node + (int64_t)S * c + s;
        *f = *f + Fk[q] * w * r[3 + c];
      }
  }
  for (int dim = 2; dim >= 0; --dim) {
    gridify_fold_n(g, N, dim, S);
    gridify_fold_n(g, F, dim, 3 * S);
  }
  for (int s = 0; s < S; ++s)
    for (int64_t q = 0; q < G; ++q) C[q] = C[q] + N[(int64_t)S * q + s] * charge[s];
  return skipped;
}
