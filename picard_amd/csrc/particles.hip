// particles.hip -- FP64 particle kernels on structure-of-arrays streams:
// AdvancePosition, BCoutflow (parallel swap-with-last), ParticlesTransfer
// (one-domain wrap), AdvanceVelocity (CIC gather + Boris), and the AoS <->
// SoA staging used by the host boundary.
//
// All arithmetic follows the reference's operation order and is compiled
// without FMA contraction, so results are bit-identical to the Fortran.
#include "cic.h"
#include "picard_internal.h"

namespace picard {
namespace {

constexpr int kT = 256;

// AdvancePosition.f90:15
__global__ void __launch_bounds__(kT) k_push_x(ParticlesF64 p, int64_t first, int64_t count,
                                                double dt) {
  const int64_t i = first + (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= first + count) return;
  p.r[0][i] = p.r[0][i] + p.r[3][i] * dt;
  p.r[1][i] = p.r[1][i] + p.r[4][i] * dt;
  p.r[2][i] = p.r[2][i] + p.r[5][i] * dt;
}

// BCoutflow.f90:28-66: the delete decision of one particle.
__device__ __forceinline__ bool outflow_delete(const GridDims &g, double x, double y, double z) {
  bool del = false;
  if (g.flow[0] && (x < g.gmin[0] || x > g.gmax[0])) del = true;
  if (g.flow[1] && (y < g.gmin[1] || y > g.gmax[1])) del = true;
  if (g.flow[2] && (z < g.gmin[2] || z > g.gmax[2])) del = true;
  if (x < 2.0 * g.lmin[0] - g.lmax[0] || x > 2.0 * g.lmax[0] - g.lmin[0]) del = true;
  if (y < 2.0 * g.lmin[1] - g.lmax[1] || y > 2.0 * g.lmax[1] - g.lmin[1]) del = true;
  if (z < 2.0 * g.lmin[2] - g.lmax[2] || z > 2.0 * g.lmax[2] - g.lmin[2]) del = true;
  return del;
}

__global__ void __launch_bounds__(kT) k_outflow_flag(ParticlesF64 p, int64_t n, GridDims g,
                                                      int32_t *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  flags[i] = outflow_delete(g, p.r[0][i], p.r[1][i], p.r[2][i]) ? 1 : 0;
}

// The sequential loop (BCoutflow.f90:24-83) runs pp = n..1 and moves the
// current last particle into each deleted slot.  Its result, in parallel
// (n' = n - |deleted|, 0-based slots):
//  * slots >= n' are visited in descending order; a kept slot enqueues its
//    particle, a deleted slot with a kept slot above it dequeues the oldest
//    pending particle and enqueues it again (that particle moved into the
//    hole and will move once more), a deleted slot with nothing above is a
//    plain pop;
//  * deleted slots < n', in descending order, take the pending particles in
//    dequeue order.
// The enqueue stream e[t] holds either a particle index or a pointer to an
// earlier stream entry (the one a rotation dequeued); pointer jumping
// resolves it, and the hole of descending rank r receives e[R + r], R being
// the number of rotations.  `scan` is the exclusive prefix count of
// deletions, scan[n] = total.

__global__ void __launch_bounds__(kT) k_outflow_rot(int64_t n, int64_t nn,
                                                     const int32_t *__restrict__ flags,
                                                     const int32_t *__restrict__ scan,
                                                     int32_t *__restrict__ rot) {
  const int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (a >= n - nn) return;
  const int64_t q = nn + a;
  const int64_t dtot = scan[n];
  const int64_t kept_above = (n - 1 - q) - (dtot - scan[q + 1]);
  rot[a] = (flags[q] && kept_above > 0) ? 1 : 0;
}

__global__ void __launch_bounds__(kT) k_outflow_stream(int64_t n, int64_t nn,
                                                        const int32_t *__restrict__ flags,
                                                        const int32_t *__restrict__ scan,
                                                        const int32_t *__restrict__ scan_rot,
                                                        int32_t *__restrict__ val,
                                                        int32_t *__restrict__ ptr) {
  const int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t W = n - nn;
  if (a >= W) return;
  const int64_t q = nn + a;
  const int64_t dtot = scan[n];
  const int64_t kept_above = (n - 1 - q) - (dtot - scan[q + 1]);
  const int64_t rtot = scan_rot[W];
  const int64_t rot_above = rtot - scan_rot[a + 1];
  const int64_t t = kept_above + rot_above;
  if (!flags[q]) {
    val[t] = (int32_t)q;
    ptr[t] = -1;
  } else if (kept_above > 0) {
    val[t] = -1;
    ptr[t] = (int32_t)rot_above;
  }
}

__global__ void __launch_bounds__(kT) k_pointer_jump(int64_t T, const int32_t *__restrict__ vin,
                                                      const int32_t *__restrict__ pin,
                                                      int32_t *__restrict__ vout,
                                                      int32_t *__restrict__ pout) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= T) return;
  const int32_t v = vin[t];
  if (v >= 0) {
    vout[t] = v;
    pout[t] = -1;
    return;
  }
  const int32_t u = pin[t];
  const int32_t vu = vin[u];
  if (vu >= 0) {
    vout[t] = vu;
    pout[t] = -1;
  } else {
    vout[t] = -1;
    pout[t] = pin[u];
  }
}

__global__ void __launch_bounds__(kT) k_outflow_fill(ParticlesF64 p, int64_t nn,
                                                      const int32_t *__restrict__ flags,
                                                      const int32_t *__restrict__ scan,
                                                      const int32_t *__restrict__ val,
                                                      int64_t rtot) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nn || !flags[i]) return;
  const int64_t m = scan[nn];
  const int64_t r = m - 1 - scan[i];
  const int64_t src = val[rtot + r];
#pragma unroll
  for (int c = 0; c < 8; ++c) p.r[c][i] = p.r[c][src];
  p.sp[i] = p.sp[src];
}

// ParticlesTransfer.f90:494-519 with idpz = idpy = idpx = myid.
__global__ void __launch_bounds__(kT) k_wrap(ParticlesF64 p, int64_t n, GridDims g) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double Lx = g.gmax[0] - g.gmin[0];
  const double Ly = g.gmax[1] - g.gmin[1];
  const double Lz = g.gmax[2] - g.gmin[2];
  double z = p.r[2][i];
  if (z < g.gmin[2]) z = z + Lz;
  if (z > g.gmax[2]) z = z - Lz;
  double y = p.r[1][i];
  if (y < g.gmin[1]) y = y + Ly;
  if (y > g.gmax[1]) y = y - Ly;
  double x = p.r[0][i];
  if (x < g.gmin[0]) x = x + Lx;
  if (x > g.gmax[0]) x = x - Lx;
  p.r[0][i] = x;
  p.r[1][i] = y;
  p.r[2][i] = z;
}

struct BorisConst {
  double etadt2[8];
  double tB[8][3], sB[8][3];
  double ve[3], B0[3];
  int32_t fluid[8], gyration[8];
};

// MODE 0: AdvanceVelocity (AdvanceVelocity.f90:34-121), the Boris push.
// MODE 1: AdvanceVelocityImp (:240-347), the electric half kick then the
//         norm-preserving rotation, leaving v at the half step of output steps.
// MODE 2: AdvanceVelocityExp (:128-238), the rotation then the electric kick.
// etadt2 and tB in `bc` are the mode's own (scaled by 2 deltat/dt for 1, 2).
// One thread per particle.
template <int MODE>
__global__ void __launch_bounds__(kT) k_boris(ParticlesF64 p, int64_t n, GridDims g,
                                               const double *__restrict__ E, BorisConst bc,
                                               int32_t *__restrict__ errors) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const CicLoc L = cic_locate(g, p.r[0][i], p.r[1][i], p.r[2][i]);
  if (!cic_in_bounds(g, L)) {
    atomicAdd(errors, 1);
    return;
  }
  const double *rs = L.rest, *ds = L.drest;
  const double F1 = ds[0] * ds[1] * ds[2];
  const double F2 = rs[0] * ds[1] * ds[2];
  const double F3 = ds[0] * rs[1] * ds[2];
  const double F4 = ds[0] * ds[1] * rs[2];
  const double F5 = rs[0] * rs[1] * ds[2];
  const double F6 = ds[0] * rs[1] * rs[2];
  const double F7 = rs[0] * ds[1] * rs[2];
  const double F8 = rs[0] * rs[1] * rs[2];
  const int64_t X = g.X, XY = (int64_t)g.X * g.Y;
  const int64_t c111 = node_index(g, L.idx[0], L.idx[1], L.idx[2]);
  const double *e111 = E + 3 * c111;
  const double *e211 = E + 3 * (c111 + 1);
  const double *e121 = E + 3 * (c111 + X);
  const double *e112 = E + 3 * (c111 + XY);
  const double *e221 = E + 3 * (c111 + 1 + X);
  const double *e122 = E + 3 * (c111 + X + XY);
  const double *e212 = E + 3 * (c111 + 1 + XY);
  const double *e222 = E + 3 * (c111 + 1 + X + XY);
  double Ep[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double a = e111[c] * F1 + e211[c] * F2;
    a = a + e121[c] * F3;
    a = a + e112[c] * F4;
    a = a + e221[c] * F5;
    a = a + e122[c] * F6;
    a = a + e212[c] * F7;
    a = a + e222[c] * F8;
    Ep[c] = a / g.dV;
  }
  const int s = p.sp[i] - 1;
  double v1 = p.r[3][i], v2 = p.r[4][i], v3 = p.r[5][i];
  if (bc.fluid[s]) {
    if (!bc.gyration[s]) return;
    const double B1 = bc.B0[0], B2 = bc.B0[1], B3 = bc.B0[2];
    const double vdb = v1 * B1 + v2 * B2 + v3 * B3;
    const double vd1 = vdb * B1 + (Ep[1] * B3 - Ep[2] * B2);
    const double vd2 = vdb * B2 + (Ep[2] * B1 - Ep[0] * B3);
    const double vd3 = vdb * B3 + (Ep[0] * B2 - Ep[1] * B1);
    const double inv = 1.0 / (B1 * B1 + B2 * B2 + B3 * B3);
    p.r[3][i] = vd1 * inv;
    p.r[4][i] = vd2 * inv;
    p.r[5][i] = vd3 * inv;
    return;
  }
  const double h = bc.etadt2[s];
  const double *t = bc.tB[s], *q = bc.sB[s];
  if (MODE == 0) {
    const double vm1 = v1 - bc.ve[0] + h * Ep[0];
    const double vm2 = v2 - bc.ve[1] + h * Ep[1];
    const double vm3 = v3 - bc.ve[2] + h * Ep[2];
    const double vd1 = vm1 + vm2 * t[2] - vm3 * t[1];
    const double vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
    const double vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
    const double vp1 = vm1 + vd2 * q[2] - vd3 * q[1];
    const double vp2 = vm2 + vd3 * q[0] - vd1 * q[2];
    const double vp3 = vm3 + vd1 * q[1] - vd2 * q[0];
    p.r[3][i] = vp1 + bc.ve[0] + h * Ep[0];
    p.r[4][i] = vp2 + bc.ve[1] + h * Ep[1];
    p.r[5][i] = vp3 + bc.ve[2] + h * Ep[2];
    return;
  }
  // Imp: kick in the ve frame before the rotation; Exp: after it
  double vm1 = v1 - bc.ve[0], vm2 = v2 - bc.ve[1], vm3 = v3 - bc.ve[2];
  if (MODE == 1) {
    vm1 = vm1 + h * Ep[0];
    vm2 = vm2 + h * Ep[1];
    vm3 = vm3 + h * Ep[2];
  }
  double vd1 = vm1 + vm2 * t[2] - vm3 * t[1];
  double vd2 = vm2 + vm3 * t[0] - vm1 * t[2];
  double vd3 = vm3 + vm1 * t[1] - vm2 * t[0];
  const double sd = vd1 * vd1 + vd2 * vd2 + vd3 * vd3;
  if (sd > 0.0) {                                   // normalise (:207-209, :333-335)
    const double f = sqrt((vm1 * vm1 + vm2 * vm2 + vm3 * vm3) / sd);
    vd1 = vd1 * f;
    vd2 = vd2 * f;
    vd3 = vd3 * f;
  }
  if (MODE == 1) {
    p.r[3][i] = vd1 + bc.ve[0];
    p.r[4][i] = vd2 + bc.ve[1];
    p.r[5][i] = vd3 + bc.ve[2];
  } else {
    p.r[3][i] = vd1 + bc.ve[0] + h * Ep[0];
    p.r[4][i] = vd2 + bc.ve[1] + h * Ep[1];
    p.r[5][i] = vd3 + bc.ve[2] + h * Ep[2];
  }
}

__global__ void __launch_bounds__(kT) k_aos_to_soa(const double *__restrict__ stage,
                                                    const int32_t *__restrict__ stage_sp,
                                                    int64_t count, ParticlesF64 p, int64_t at) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= count) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) p.r[c][at + i] = stage[8 * i + c];
  p.sp[at + i] = stage_sp[i];
}

__global__ void __launch_bounds__(kT) k_soa_to_aos(ParticlesF64 p, int64_t first, int64_t count,
                                                    double *__restrict__ stage,
                                                    int32_t *__restrict__ stage_sp) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= count) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) stage[8 * i + c] = p.r[c][first + i];
  stage_sp[i] = p.sp[first + i];
}

constexpr int64_t kStageChunk = int64_t(1) << 21;

}  // namespace

int ensure_stage(picard_ctx *c, int64_t want) {
  if (c->s.stage_cap >= want) return PICARD_OK;
  if (c->s.stage) PICARD_HIP(hipFree(c->s.stage));
  if (c->s.stage_sp) PICARD_HIP(hipFree(c->s.stage_sp));
  c->s.stage = nullptr;
  c->s.stage_sp = nullptr;
  PICARD_HIP(hipMalloc(&c->s.stage, sizeof(double) * 8 * want));
  PICARD_HIP(hipMalloc(&c->s.stage_sp, sizeof(int32_t) * want));
  c->s.stage_cap = want;
  return PICARD_OK;
}

int launch_push_x(picard_ctx *c, int64_t first, int64_t count, double dt) {
  if (count <= 0) return PICARD_OK;
  k_push_x<<<grid_blocks(count, kT), kT, 0, c->stream>>>(c->p, first, count, dt);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int launch_outflow(picard_ctx *c, int64_t *n_new, int64_t *removed) {
  const int64_t n = c->n;
  *n_new = n;
  *removed = 0;
  if (n == 0) return PICARD_OK;
  int32_t *flags = c->s.flags, *scan = c->s.scan;
  k_outflow_flag<<<grid_blocks(n, kT), kT, 0, c->stream>>>(c->p, n, c->gd, flags);
  PICARD_HIP(hipGetLastError());
  PICARD_TRY(scan_exclusive_i32(c, flags, scan, n));
  PICARD_HIP(hipMemcpyAsync(c->host_counters, scan + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  const int64_t dtot = c->host_counters[0];
  if (dtot == 0) return PICARD_OK;
  const int64_t nn = n - dtot;
  const int64_t W = dtot;                       // slots above n'
  int32_t *rot = c->s.work[0], *scan_rot = c->s.work[1];
  int32_t *val[2] = {c->s.work[2], c->s.work[3]}, *ptr[2] = {c->s.work[4], c->s.work[5]};
  k_outflow_rot<<<grid_blocks(W, kT), kT, 0, c->stream>>>(n, nn, flags, scan, rot);
  PICARD_HIP(hipGetLastError());
  PICARD_TRY(scan_exclusive_i32(c, rot, scan_rot, W));
  k_outflow_stream<<<grid_blocks(W, kT), kT, 0, c->stream>>>(n, nn, flags, scan, scan_rot,
                                                             val[0], ptr[0]);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipMemcpyAsync(c->host_counters, scan_rot + W, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipMemcpyAsync(c->host_counters + 1, scan + nn, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  const int64_t rtot = c->host_counters[0], m = c->host_counters[1];
  const int64_t T = m + rtot;
  int cur = 0;
  if (rtot > 0) {
    int rounds = 1;
    while ((int64_t(1) << rounds) < T) ++rounds;
    for (int q = 0; q <= rounds; ++q) {
      k_pointer_jump<<<grid_blocks(T, kT), kT, 0, c->stream>>>(T, val[cur], ptr[cur],
                                                               val[cur ^ 1], ptr[cur ^ 1]);
      PICARD_HIP(hipGetLastError());
      cur ^= 1;
    }
  }
  if (m > 0) {
    k_outflow_fill<<<grid_blocks(nn, kT), kT, 0, c->stream>>>(c->p, nn, flags, scan, val[cur],
                                                              rtot);
    PICARD_HIP(hipGetLastError());
  }
  c->n = nn;
  *n_new = nn;
  *removed = dtot;
  return PICARD_OK;
}

int launch_wrap(picard_ctx *c) {
  if (c->n <= 0) return PICARD_OK;
  k_wrap<<<grid_blocks(c->n, kT), kT, 0, c->stream>>>(c->p, c->n, c->gd);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int launch_boris(picard_ctx *c, double dt, int mode, double deltat) {
  if (c->n <= 0) return PICARD_OK;
  BorisConst bc;
  for (int s = 0; s < 8; ++s) {
    if (mode == 0) {
      bc.etadt2[s] = 0.5 * dt * c->spc.eta[s];               // AdvanceVelocity.f90:35
      for (int q = 0; q < 3; ++q) bc.tB[s][q] = c->spc.tB[4 * s + q];
    } else {                                                // :156-157, :267-268
      bc.etadt2[s] = 0.5 * dt * c->spc.eta[s] * 2.0 * deltat / dt;
      for (int q = 0; q < 3; ++q) bc.tB[s][q] = c->spc.tB[4 * s + q] * 2.0 * deltat / dt;
    }
    for (int q = 0; q < 3; ++q) bc.sB[s][q] = c->spc.sB[4 * s + q];
    bc.fluid[s] = c->spc.fluid[s];
    bc.gyration[s] = c->spc.gyration[s];
  }
  for (int q = 0; q < 3; ++q) {
    bc.ve[q] = c->spc.ve[q];
    bc.B0[q] = c->spc.B0[q];
  }
  PICARD_HIP(hipMemsetAsync(c->s.counters, 0, sizeof(int32_t), c->stream));
  const int nb = grid_blocks(c->n, kT);
  if (mode == 0)
    k_boris<0><<<nb, kT, 0, c->stream>>>(c->p, c->n, c->gd, c->E, bc, c->s.counters);
  else if (mode == 1)
    k_boris<1><<<nb, kT, 0, c->stream>>>(c->p, c->n, c->gd, c->E, bc, c->s.counters);
  else
    k_boris<2><<<nb, kT, 0, c->stream>>>(c->p, c->n, c->gd, c->E, bc, c->s.counters);
  PICARD_HIP(hipGetLastError());
  PICARD_HIP(hipMemcpyAsync(c->host_counters, c->s.counters, sizeof(int32_t),
                            hipMemcpyDeviceToHost, c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  if (c->host_counters[0] != 0) {
    set_error("AdvanceVelocity: " + std::to_string(c->host_counters[0]) +
              " particle(s) outside the field arrays were left unchanged");
    return PICARD_ERR_RANGE;
  }
  return PICARD_OK;
}

int launch_append(picard_ctx *c, const double *host_rp, const int32_t *host_sp, int64_t count) {
  if (count <= 0) return PICARD_OK;
  if (c->n + count > c->cap) {
    set_error("particles_append: capacity max_particles exceeded");
    return PICARD_ERR_CAPACITY;
  }
  PICARD_TRY(ensure_stage(c, count < kStageChunk ? count : kStageChunk));
  for (int64_t done = 0; done < count; done += kStageChunk) {
    const int64_t m = (count - done) < kStageChunk ? (count - done) : kStageChunk;
    PICARD_HIP(hipMemcpyAsync(c->s.stage, host_rp + 8 * done, sizeof(double) * 8 * m,
                              hipMemcpyHostToDevice, c->stream));
    PICARD_HIP(hipMemcpyAsync(c->s.stage_sp, host_sp + done, sizeof(int32_t) * m,
                              hipMemcpyHostToDevice, c->stream));
    k_aos_to_soa<<<grid_blocks(m, kT), kT, 0, c->stream>>>(c->s.stage, c->s.stage_sp, m, c->p,
                                                           c->n + done);
    PICARD_HIP(hipGetLastError());
    PICARD_HIP(hipStreamSynchronize(c->stream));
  }
  c->n += count;
  return PICARD_OK;
}

int stage_particles(picard_ctx *c, int64_t first, int64_t count, double *stage, int32_t *stage_sp) {
  if (count <= 0) return PICARD_OK;
  if (c->split) return sp_stage(c, first, count, stage, stage_sp);
  k_soa_to_aos<<<grid_blocks(count, kT), kT, 0, c->stream>>>(c->p, first, count, stage, stage_sp);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int launch_read(picard_ctx *c, int64_t first, int64_t count, double *host_rp, int32_t *host_sp) {
  if (count <= 0) return PICARD_OK;
  if (first < 0 || first + count > c->n) {
    set_error("particles_read: range outside [0, num_local)");
    return PICARD_ERR_ARG;
  }
  PICARD_TRY(ensure_stage(c, count < kStageChunk ? count : kStageChunk));
  for (int64_t done = 0; done < count; done += kStageChunk) {
    const int64_t m = (count - done) < kStageChunk ? (count - done) : kStageChunk;
    k_soa_to_aos<<<grid_blocks(m, kT), kT, 0, c->stream>>>(c->p, first + done, m, c->s.stage,
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

}  // namespace picard
