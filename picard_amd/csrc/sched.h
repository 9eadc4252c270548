// sched.h -- XCD-aware persistent schedule of the interior cells for the
// stencil kernels (Jacobi sweeps, residuals, updates).
//
// The interior is cut into tiles of bx x by cells of one z-plane (bx*by =
// 256 threads, x fastest).  The launch has 8*W blocks; workgroups are
// dispatched to the 8 XCDs round robin, so block b runs on XCD b % 8.  XCD x
// owns the z-planes [x*nz/8, (x+1)*nz/8) and its W blocks march through them
// plane by plane, W tiles of a plane at a time: the z-neighbours a sweep reads
// were touched by the same XCD a plane earlier and are still in its L2.  The
// assignment of cells to blocks is fixed, so per-block partial sums (and the
// fixed-order sum over them) are deterministic.  If the dispatch order differs
// the schedule stays correct, only less cache friendly.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace picard {

struct TileSched {
  int bx, by;        // tile shape (threads per block = bx * by)
  int ntx, nty;      // tiles per plane along x, y
  int TP;            // tiles per plane
  int W;             // blocks per XCD
  int B;             // blocks in the launch (8 * W)
};

inline TileSched make_sched(int nx, int ny, int nz, int threads = 256) {
  TileSched s;
  s.bx = 64;
  while (s.bx > 8 && s.bx / 2 >= nx) s.bx /= 2;
  s.by = threads / s.bx;
  s.ntx = (nx + s.bx - 1) / s.bx;
  s.nty = (ny + s.by - 1) / s.by;
  s.TP = s.ntx * s.nty;
  // enough blocks per XCD to cover a plane, at most 256 (8 waves per CU of
  // its 32 CUs), at least 4
  s.W = s.TP < 256 ? s.TP : 256;
  if (s.W < 4) s.W = 4;
  s.B = 8 * s.W;
  (void)nz;
  return s;
}

// Visit every (i, j, k) interior cell (1-based Fortran node indices) this
// block owns; f(i, j, k, valid).  All threads of the block take the same
// number of iterations, so the body may use __syncthreads.
template <typename F>
__device__ __forceinline__ void for_my_cells(const TileSched &s, int nx, int ny, int nz, F &&f) {
  const int b = blockIdx.x;
  const int x = b & 7, w = b >> 3;
  const int z0 = (x * nz) >> 3, z1 = ((x + 1) * nz) >> 3;
  for (int z = z0; z < z1; ++z) {
    for (int t = w; t < s.TP; t += s.W) {
      const int tx = t % s.ntx, ty = t / s.ntx;
      const int i = tx * s.bx + (int)threadIdx.x + 2;
      const int j = ty * s.by + (int)threadIdx.y + 2;
      f(i, j, z + 2, i <= nx + 1 && j <= ny + 1);
    }
  }
}

}  // namespace picard
