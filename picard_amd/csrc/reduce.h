// reduce.h -- deterministic block and grid reductions for the stencil kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace picard {

__device__ __forceinline__ int lin_tid() {
  return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ unsigned lin_bid() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ unsigned n_blocks() { return gridDim.x * gridDim.y * gridDim.z; }

// Sum over the block (at most 1024 threads; `sh` holds one double per
// wave): wave64 shuffle tree, then the waves in order.  Result in thread 0;
// every thread must call it.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int t = lin_tid();
  const int lane = t & 63, wave = t >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0) {
    const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
    for (int w = 0; w < nw; ++w) s = s + sh[w];
  }
  return s;
}

// Sum of the per-block partials by the whole (last) block: each thread adds a
// fixed strided subset, then the fixed block tree; the result (thread 0)
// does not depend on which block finished last.
__device__ __forceinline__ double final_sum(const double *partials, unsigned n, double *sh) {
  __threadfence();
  double s = 0.0;
  const unsigned nt = blockDim.x * blockDim.y * blockDim.z;
  for (unsigned q = lin_tid(); q < n; q += nt) s = s + partials[q];
  return block_sum(s, sh);
}

// Kernel duration of a sampled launch: each block's first thread takes the
// GPU's constant 100 MHz real-time clock when the block starts and when it
// has finished.  The earliest start and the latest end are kept in
// kStampWays words each, a block using word blockIdx.x % kStampWays, so that
// few blocks meet on one address (one word per launch held its blocks up by
// about 2 us); ks[0 .. kStampWays) keep starts (reset to ~0), ks[kStampWays
// .. 2 kStampWays) ends (reset to 0), and the host reduces them.  ks is null
// in launches not sampled.
constexpr int kStampWays = 16;
__device__ __forceinline__ void kstamp_begin(unsigned long long *ks) {
  if (ks && lin_tid() == 0)
    atomicMin(ks + (blockIdx.x % kStampWays), (unsigned long long)__builtin_amdgcn_s_memrealtime());
}
// Every thread of the block calls it (block-uniform), after the block's work.
__device__ __forceinline__ void kstamp_end(unsigned long long *ks) {
  if (!ks) return;
  __syncthreads();
  if (lin_tid() == 0)
    atomicMax(ks + kStampWays + (blockIdx.x % kStampWays),
              (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

}  // namespace picard
