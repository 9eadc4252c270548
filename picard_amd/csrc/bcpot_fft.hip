// bcpot_fft.hip -- BCpotentials (BCpotentials.f90:51-141) for the FP32-split
// form: the open-boundary face sums in FP64, evaluated as FFT convolutions.
//
// For a flow dimension d, face f and the plane n of cells along d, the
// reference sums C(cell) * D(offset) over the plane's cells, and the offset
// across the plane is (t1 - tt1, t2 - tt2) for face node (tt1, tt2): a 2D
// convolution of the charge plane with an even kernel K_{f,n}(o1, o2) taken
// from the distance table.  So
//     U_f = IFFT( sum_n FFT(C_n) * FFT(K_{f,n}) ),
// on an M1 x M2 grid (powers of two >= 2 N_t, so the circular convolution
// has no aliasing on the face nodes).  FFT(K_{f,n}) depends on the geometry
// only and is computed once.  Summation order differs from the reference's
// sequential sum; the FP64 form keeps the sequential kernel (grid.hip).
//
// The FFT is a radix-2 Cooley-Tukey transform in LDS, one workgroup per
// transform of up to 1024 points, batched over rows and columns.
#include "picard_internal.h"

namespace picard {
namespace {

constexpr int kT = 256;

// In-LDS FFT of `L` points (power of two <= 1024) for the batch of
// transforms x[y*pstride + b*tstride + e*estride], e = 0..L-1, with
// (b, y) = (blockIdx.x, blockIdx.y).  sign = -1 forward, +1
// inverse (unnormalised).  blockDim.x = min(L/2, 512).
__global__ void k_fft(double2 *__restrict__ x, int L, int logL, int64_t tstride, int64_t estride,
                      int64_t pstride, const double2 *__restrict__ tw, int inverse) {
  __shared__ double2 sm[1024];
  const int64_t base = (int64_t)blockIdx.y * pstride + (int64_t)blockIdx.x * tstride;
  // bit-reversed load
  for (int e = threadIdx.x; e < L; e += blockDim.x) {
    const int r = __brev((unsigned)e) >> (32 - logL);
    sm[r] = x[base + (int64_t)e * estride];
  }
  __syncthreads();
  for (int s = 1; s <= logL; ++s) {
    const int half = 1 << (s - 1);
    const int step = L >> s;   // twiddle stride in the table of L/2 entries
    for (int t = threadIdx.x; t < L / 2; t += blockDim.x) {
      const int grp = t >> (s - 1), pos = t & (half - 1);
      const int i0 = grp * (2 * half) + pos, i1 = i0 + half;
      double2 w = tw[pos * step];
      if (inverse) w.y = -w.y;
      const double2 a = sm[i0], b = sm[i1];
      const double2 bw = make_double2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
      sm[i0] = make_double2(a.x + bw.x, a.y + bw.y);
      sm[i1] = make_double2(a.x - bw.x, a.y - bw.y);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < L; e += blockDim.x) x[base + (int64_t)e * estride] = sm[e];
}

__global__ void k_twiddles(double2 *tw, int L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L / 2) return;
  double s, c;
  sincospi(-2.0 * (double)k / (double)L, &s, &c);
  tw[k] = make_double2(c, s);
}

// Geometry of one flow dimension: normal dim d, transverse dims t1 < t2.
struct FaceGeo {
  int d, t1, t2;
  int nd, n1, n2;        // interior cells along d, t1, t2
  int M1, M2, log1, log2;
  int64_t sd, s1, s2;    // node strides of d, t1, t2 in U / C
  int64_t dsd, ds1, ds2; // strides of d, t1, t2 in the distance table
  int cd, c1, c2;        // distance-table centres (N + 1, 0-based N)
};

// Kernel of face f, plane n (1-based cell index along d) on the padded grid:
// K(o1, o2) = D(normal offset, |o1|, |o2|), zero outside |o| <= n - 1.
__global__ void __launch_bounds__(kT) k_fill_kernel(FaceGeo fg, const double *__restrict__ D,
                                                    double2 *__restrict__ K, int face) {
  const int64_t plane_sz = (int64_t)fg.M1 * fg.M2;
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= plane_sz * fg.nd) return;
  const int n = (int)(t / plane_sz) + 1;
  const int64_t r = t % plane_sz;
  const int o1 = (int)(r / fg.M2), o2 = (int)(r % fg.M2);
  const int off1 = o1 < fg.M1 / 2 ? o1 : o1 - fg.M1;
  const int off2 = o2 < fg.M2 / 2 ? o2 : o2 - fg.M2;
  double v = 0.0;
  if (off1 > -fg.n1 && off1 < fg.n1 && off2 > -fg.n2 && off2 < fg.n2) {
    const int offd = face == 0 ? n : n - fg.nd - 1;
    const int64_t idx = (int64_t)(fg.cd + offd) * fg.dsd + (int64_t)(fg.c1 + off1) * fg.ds1 +
                        (int64_t)(fg.c2 + off2) * fg.ds2;
    v = D[idx];
  }
  K[t] = make_double2(v, 0.0);
}

// Zero-padded charge planes: P[n][o1][o2] = C(cell n, o1, o2).
__global__ void __launch_bounds__(kT) k_pack_planes(FaceGeo fg, const double *__restrict__ C,
                                                    double2 *__restrict__ P) {
  const int64_t plane_sz = (int64_t)fg.M1 * fg.M2;
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= plane_sz * fg.nd) return;
  const int n = (int)(t / plane_sz);
  const int64_t r = t % plane_sz;
  const int o1 = (int)(r / fg.M2), o2 = (int)(r % fg.M2);
  double v = 0.0;
  if (o1 < fg.n1 && o2 < fg.n2)
    v = C[(int64_t)(n + 1) * fg.sd + (int64_t)(o1 + 1) * fg.s1 + (int64_t)(o2 + 1) * fg.s2];
  P[t] = make_double2(v, 0.0);
}

// Uhat_f = sum_n P_n * Khat_{f,n} for both faces.
__global__ void __launch_bounds__(kT) k_accumulate(FaceGeo fg, const double2 *__restrict__ P,
                                                   const double2 *__restrict__ K0,
                                                   const double2 *__restrict__ K1,
                                                   double2 *__restrict__ U0, double2 *__restrict__ U1) {
  const int64_t plane_sz = (int64_t)fg.M1 * fg.M2;
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= plane_sz) return;
  double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
  for (int n = 0; n < fg.nd; ++n) {
    const double2 p = P[(int64_t)n * plane_sz + r];
    const double2 k0 = K0[(int64_t)n * plane_sz + r];
    const double2 k1 = K1[(int64_t)n * plane_sz + r];
    a0r += p.x * k0.x - p.y * k0.y;
    a0i += p.x * k0.y + p.y * k0.x;
    a1r += p.x * k1.x - p.y * k1.y;
    a1i += p.x * k1.y + p.y * k1.x;
  }
  U0[r] = make_double2(a0r, a0i);
  U1[r] = make_double2(a1r, a1i);
}

__global__ void __launch_bounds__(kT) k_write_faces(FaceGeo fg, const double2 *__restrict__ U0,
                                                    const double2 *__restrict__ U1,
                                                    double *__restrict__ U, double scale) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)fg.n1 * fg.n2) return;
  const int a = (int)(t % fg.n1), b = (int)(t / fg.n1);
  const int64_t r = (int64_t)a * fg.M2 + b;
  const int64_t at = (int64_t)(a + 1) * fg.s1 + (int64_t)(b + 1) * fg.s2;
  U[at] = U0[r].x * scale;                                     // guard plane 1
  U[at + (int64_t)(fg.nd + 1) * fg.sd] = U1[r].x * scale;      // guard plane N + 2
}

int pow2_at_least(int v, int *lg) {
  int m = 1, l = 0;
  while (m < v) {
    m <<= 1;
    ++l;
  }
  *lg = l;
  return m;
}

FaceGeo make_face_geo(const GridDims &g, int d) {
  FaceGeo fg;
  const int n[3] = {g.nx, g.ny, g.nz};
  const int64_t stride[3] = {1, (int64_t)g.X, (int64_t)g.X * g.Y};
  const int64_t DX = 2 * (int64_t)g.nx + 1, DY = 2 * (int64_t)g.ny + 1;
  const int64_t dstride[3] = {1, DX, DX * DY};
  fg.d = d;
  fg.t1 = d == 0 ? 1 : 0;
  fg.t2 = d == 2 ? 1 : 2;
  fg.nd = n[d];
  fg.n1 = n[fg.t1];
  fg.n2 = n[fg.t2];
  fg.M1 = pow2_at_least(2 * fg.n1, &fg.log1);
  fg.M2 = pow2_at_least(2 * fg.n2, &fg.log2);
  fg.sd = stride[d];
  fg.s1 = stride[fg.t1];
  fg.s2 = stride[fg.t2];
  fg.dsd = dstride[d];
  fg.ds1 = dstride[fg.t1];
  fg.ds2 = dstride[fg.t2];
  fg.cd = n[d];         // 0-based index of the table centre N + 1
  fg.c1 = n[fg.t1];
  fg.c2 = n[fg.t2];
  return fg;
}

inline int fft_threads(int L) {
  const int h = L / 2;
  return h < 64 ? 64 : (h > 512 ? 512 : h);
}

// 2D FFT of `batch` planes of M1 x M2 (row-major, o2 fastest).
int fft2d(picard_ctx *c, double2 *x, int64_t batch, const FaceGeo &fg, double2 *const tw[2],
          int inverse) {
  const int64_t plane = (int64_t)fg.M1 * fg.M2;
  // rows: batch*M1 contiguous transforms of length M2
  k_fft<<<dim3((unsigned)fg.M1, (unsigned)batch), fft_threads(fg.M2), 0, c->stream>>>(
      x, fg.M2, fg.log2, fg.M2, 1, plane, tw[1], inverse);
  PICARD_HIP(hipGetLastError());
  // columns: M2 transforms of length M1 (element stride M2) per plane
  k_fft<<<dim3((unsigned)fg.M2, (unsigned)batch), fft_threads(fg.M1), 0, c->stream>>>(
      x, fg.M1, fg.log1, 1, fg.M2, plane, tw[0], inverse);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

}  // namespace

void bcpot_fft_release(picard_ctx *c) {
  for (int d = 0; d < 3; ++d)
    for (int f = 0; f < 2; ++f) {
      if (c->fft_K[d][f]) (void)hipFree(c->fft_K[d][f]);
      if (c->fft_tw[d][f]) (void)hipFree(c->fft_tw[d][f]);
      c->fft_K[d][f] = nullptr;
      c->fft_tw[d][f] = nullptr;
    }
  if (c->fft_work) (void)hipFree(c->fft_work);
  c->fft_work = nullptr;
  c->fft_work_len = 0;
  c->fft_ready = false;
}

// Transforms of the face kernels, FFT(K_{f,n}) for every flow dimension,
// face and plane, from the current distance table.
int bcpot_fft_prepare(picard_ctx *c) {
  const GridDims &g = c->gd;
  for (int d = 0; d < 3; ++d) {
    if (!g.flow[d]) continue;
    const FaceGeo fg = make_face_geo(g, d);
    if (fg.M1 > 1024 || fg.M2 > 1024) {
      set_error("BCpotentials (split form): faces are limited to 512 x 512 cells");
      return PICARD_ERR_ARG;
    }
    const int64_t plane = (int64_t)fg.M1 * fg.M2;
    for (int f = 0; f < 2; ++f) {
      if (!c->fft_K[d][f])
        PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&c->fft_K[d][f]),
                             sizeof(double2) * plane * fg.nd));
    }
    for (int a = 0; a < 2; ++a) {
      const int L = a == 0 ? fg.M1 : fg.M2;
      if (!c->fft_tw[d][a])
        PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&c->fft_tw[d][a]), sizeof(double2) * 512));
      k_twiddles<<<grid_blocks(L / 2, 256), 256, 0, c->stream>>>(c->fft_tw[d][a], L);
      PICARD_HIP(hipGetLastError());
    }
    // work: nd transformed charge planes + the two face accumulators
    if (c->fft_work_len < plane * (fg.nd + 2)) {
      if (c->fft_work) PICARD_HIP(hipFree(c->fft_work));
      c->fft_work = nullptr;
      PICARD_HIP(hipMalloc(reinterpret_cast<void **>(&c->fft_work),
                           sizeof(double2) * plane * (fg.nd + 2)));
      c->fft_work_len = plane * (fg.nd + 2);
    }
    for (int f = 0; f < 2; ++f) {
      k_fill_kernel<<<grid_blocks(plane * fg.nd, kT), kT, 0, c->stream>>>(fg, c->D,
                                                                          c->fft_K[d][f], f);
      PICARD_HIP(hipGetLastError());
      PICARD_TRY(fft2d(c, c->fft_K[d][f], fg.nd, fg, c->fft_tw[d], 0));
    }
  }
  PICARD_HIP(hipStreamSynchronize(c->stream));
  c->fft_ready = true;
  return PICARD_OK;
}

// BCpotentials.f90:36-141 on one rank: U = 0, then the potential of every
// flow face node from the interior charge.
int launch_bc_potentials_fft(picard_ctx *c) {
  const GridDims &g = c->gd;
  if (!c->have_D) {
    set_error("BCpotentials: the distance table is not set (call CalcDistance first)");
    return PICARD_ERR_STATE;
  }
  if (!c->fft_ready) PICARD_TRY(bcpot_fft_prepare(c));
  PICARD_HIP(hipMemsetAsync(c->U, 0, sizeof(double) * g.G, c->stream));   // BCpotentials.f90:36
  const double tau = 6.2831853071795864769252867665590;
  const double kfac = 1.0 / (2.0 * tau * 8.854187817e-12);
  for (int d = 0; d < 3; ++d) {
    if (!g.flow[d]) continue;
    const FaceGeo fg = make_face_geo(g, d);
    const int64_t plane = (int64_t)fg.M1 * fg.M2;
    double2 *P = c->fft_work, *U0 = P + plane * fg.nd, *U1 = U0 + plane;
    k_pack_planes<<<grid_blocks(plane * fg.nd, kT), kT, 0, c->stream>>>(fg, c->C, P);
    PICARD_HIP(hipGetLastError());
    PICARD_TRY(fft2d(c, P, fg.nd, fg, c->fft_tw[d], 0));
    k_accumulate<<<grid_blocks(plane, kT), kT, 0, c->stream>>>(fg, P, c->fft_K[d][0],
                                                               c->fft_K[d][1], U0, U1);
    PICARD_HIP(hipGetLastError());
    PICARD_TRY(fft2d(c, U0, 2, fg, c->fft_tw[d], 1));
    k_write_faces<<<grid_blocks((int64_t)fg.n1 * fg.n2, kT), kT, 0, c->stream>>>(
        fg, U0, U1, c->U, kfac / (double)plane);
    PICARD_HIP(hipGetLastError());
  }
  return PICARD_OK;
}

}  // namespace picard
