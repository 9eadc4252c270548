// diag.hip -- the diagnostics line of the main loop (picard.f90:649-672) on
// the device: particle kinetic energy, electric and magnetic field energy and
// the particle count, without downloading particles or fields.
//
//   uP = 0.5 * sum_p n2p(s) mass(s) |v|^2                   (:658-661)
//   uE = 0.5 eps0 dV * sum_interior |E + E0|^2              (:655-657)
//   uB = 0.5 / mu0 dV * sum_interior |B|^2, B = B0          (:653, :420-422)
//
// Sums run as fixed-assignment block partials followed by one fixed-order
// block reduction, so repeated calls give identical results.
#include "cic.h"
#include "picard_internal.h"
#include "reduce.h"

namespace picard {
namespace {

constexpr int kT = 256;
constexpr int kDiagBlocks = 1024;

struct DiagConst {
  double w[8];      // n2p(s) * mass(s)
  double E0[3];
};

template <typename Vel>
__global__ void __launch_bounds__(kT) k_diag_particles(Vel vel, int64_t n, DiagConst dc,
                                                       double *__restrict__ partials) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < n; q += (int64_t)gridDim.x * kT) {
    double v1, v2, v3;
    int s;
    vel(q, v1, v2, v3, s);
    acc = acc + dc.w[s] * (v1 * v1 + v2 * v2 + v3 * v3);
  }
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

template <typename Field>
__global__ void __launch_bounds__(kT) k_diag_field(GridDims g, Field ef, DiagConst dc,
                                                   double *__restrict__ partials) {
  __shared__ double sh[4];
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  double acc = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < ncell;
       q += (int64_t)gridDim.x * kT) {
    const int i = (int)(q % g.nx) + 2;
    const int j = (int)((q / g.nx) % g.ny) + 2;
    const int k = (int)(q / ((int64_t)g.nx * g.ny)) + 2;
    double e1, e2, e3;
    ef(node_index(g, i, j, k), e1, e2, e3);
    const double a = e1 + dc.E0[0], b = e2 + dc.E0[1], c = e3 + dc.E0[2];
    acc = acc + (a * a + b * b + c * c);
  }
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// out[0], out[1] = the fixed-order sums of the two partial arrays
__global__ void __launch_bounds__(kT) k_diag_final(const double *__restrict__ pp,
                                                   const double *__restrict__ pf, int nb,
                                                   double *__restrict__ out) {
  __shared__ double sh[4];
  const double a = final_sum(pp, nb, sh);
  if (threadIdx.x == 0) out[0] = a;
  __syncthreads();
  const double b = final_sum(pf, nb, sh);
  if (threadIdx.x == 0) out[1] = b;
}

struct VelF64 {
  ParticlesF64 p;
  __device__ void operator()(int64_t q, double &v1, double &v2, double &v3, int &s) const {
    v1 = p.r[3][q];
    v2 = p.r[4][q];
    v3 = p.r[5][q];
    s = p.sp[q] - 1;
  }
};

struct VelSplit {
  ParticlesSplit p;
  __device__ void operator()(int64_t q, double &v1, double &v2, double &v3, int &s) const {
    v1 = (double)p.v[0][q];
    v2 = (double)p.v[1][q];
    v3 = (double)p.v[2][q];
    s = (int)p.sp[q] - 1;
  }
};

struct FieldF64 {
  const double *E;
  __device__ void operator()(int64_t node, double &e1, double &e2, double &e3) const {
    e1 = E[3 * node];
    e2 = E[3 * node + 1];
    e3 = E[3 * node + 2];
  }
};

struct FieldF32 {
  const float *E32;   // float4 per node
  __device__ void operator()(int64_t node, double &e1, double &e2, double &e3) const {
    e1 = (double)E32[4 * node];
    e2 = (double)E32[4 * node + 1];
    e3 = (double)E32[4 * node + 2];
  }
};

}  // namespace

int diagnostics(picard_ctx *c, const double *mass, picard_diag *out) {
  const GridDims &g = c->gd;
  DiagConst dc{};
  for (int s = 0; s < 8; ++s) dc.w[s] = c->spc.n2p[s] * mass[s];
  for (int q = 0; q < 3; ++q) dc.E0[q] = c->spc.E0[q];
  const int nb = (int)(c->s.partials_len < kDiagBlocks ? c->s.partials_len : kDiagBlocks);
  double *pp = c->s.partials, *pf = c->s.partials + c->s.partials_len;
  double *res = c->s.partials + 2 * c->s.partials_len;     // two doubles
  if (c->split) {
    k_diag_particles<<<nb, kT, 0, c->stream>>>(VelSplit{c->ps}, c->n, dc, pp);
    k_diag_field<<<nb, kT, 0, c->stream>>>(g, FieldF32{c->E32}, dc, pf);
  } else {
    k_diag_particles<<<nb, kT, 0, c->stream>>>(VelF64{c->p}, c->n, dc, pp);
    k_diag_field<<<nb, kT, 0, c->stream>>>(g, FieldF64{c->E}, dc, pf);
  }
  k_diag_final<<<1, kT, 0, c->stream>>>(pp, pf, nb, res);
  PICARD_HIP(hipGetLastError());
  double h[2];
  PICARD_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  const double eps0 = 8.854187817e-12, mu0 = 1.25663706143591729538505735331180e-6;
  const double dV = g.dxyz[0] * g.dxyz[1] * g.dxyz[2];
  const double B2 = c->spc.B0[0] * c->spc.B0[0] + c->spc.B0[1] * c->spc.B0[1] +
                    c->spc.B0[2] * c->spc.B0[2];
  out->num_local = c->n;
  out->num_global = c->n;                       // MPI_reduce over one rank (:651)
  out->particle_energy = h[0] * 0.5;
  out->e_energy = h[1] * 0.5 * eps0 * dV;
  out->b_energy = B2 * (double)((int64_t)g.nx * g.ny * g.nz) * 0.5 / mu0 * dV;
  return PICARD_OK;
}

}  // namespace picard
