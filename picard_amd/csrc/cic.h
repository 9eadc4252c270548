// cic.h -- cloud-in-cell location of one particle, shared by the deposit and
// the field gather.  ParticlesGridify.f90:70-77 and AdvanceVelocity.f90:44-51
// compute the in-cell offset and the lower node index independently:
//   rest = dmod(x - Lmin + dx/2, dx)        (exact remainder, as C fmod)
//   idx  = 1 + idint((x - Lmin)/dx + 1/2)   (truncation)
// and this helper does the same, operation for operation.
#pragma once
#include "picard_internal.h"

namespace picard {

struct CicLoc {
  double rest[3], drest[3];
  int idx[3];     // 1-based lower node index
};

__device__ __forceinline__ CicLoc cic_locate(const GridDims &g, double x, double y, double z) {
  CicLoc L;
  const double r[3] = {x, y, z};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    L.rest[c] = fmod(r[c] - g.lmin[c] + 0.5 * g.dxyz[c], g.dxyz[c]);
    L.drest[c] = g.dxyz[c] - L.rest[c];
    L.idx[c] = 1 + (int)((r[c] - g.lmin[c]) / g.dxyz[c] + 0.5);
  }
  return L;
}

__device__ __forceinline__ bool cic_in_bounds(const GridDims &g, const CicLoc &L) {
  return L.idx[0] >= 1 && L.idx[1] >= 1 && L.idx[2] >= 1 && L.idx[0] <= g.nx + 1 &&
         L.idx[1] <= g.ny + 1 && L.idx[2] <= g.nz + 1;
}

// 0-based linear index of 1-based Fortran node (i, j, k)
__device__ __forceinline__ int64_t node_index(const GridDims &g, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)g.X * ((int64_t)(j - 1) + (int64_t)g.Y * (int64_t)(k - 1));
}

}  // namespace picard
