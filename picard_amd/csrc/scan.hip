// scan.hip -- device-wide exclusive prefix sum of int32 counts.
//
// Three-phase, recursive: block totals -> scan of the totals -> per-block
// scan with the block's offset.  2048 items per 256-thread block; every level
// keeps its own slice of the context's scan_tmp arena.
#include "picard_internal.h"

namespace picard {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kPerBlock = kThreads * kItems;

__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t *total) {
  __shared__ int32_t sh[kThreads];
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    int32_t add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int32_t incl = sh[t];
  *total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kThreads) k_block_totals(const int32_t *__restrict__ in, int64_t n,
                                                            int32_t *__restrict__ totals) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock;
  int32_t s = 0;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const int64_t i = base + (int64_t)q * kThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  int32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kThreads) k_block_scan(const int32_t *in, int64_t n, int32_t *out,
                                                          const int32_t *__restrict__ offsets) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock;
  // each thread owns kItems consecutive elements
  int32_t v[kItems];
  int32_t s = 0;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const int64_t i = base + (int64_t)threadIdx.x * kItems + q;
    v[q] = i < n ? in[i] : 0;
    s += v[q];
  }
  int32_t tot;
  int32_t run = block_exclusive_scan(s, &tot);
  const int32_t off = offsets ? offsets[blockIdx.x] : 0;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const int64_t i = base + (int64_t)threadIdx.x * kItems + q;
    if (i < n) out[i] = off + run;
    run += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
}

}  // namespace

int64_t scan_tmp_required(int64_t n) {
  int64_t need = 0;
  while (n > kPerBlock) {
    const int64_t nb = (n + kPerBlock - 1) / kPerBlock;
    need += nb + 1;
    n = nb;
  }
  return need + 1;
}

static int scan_level(hipStream_t st, const int32_t *in, int32_t *out, int64_t n, int32_t *tmp) {
  if (n <= kPerBlock) {
    k_block_scan<<<1, kThreads, 0, st>>>(in, n, out, nullptr);
    PICARD_HIP(hipGetLastError());
    return PICARD_OK;
  }
  const int64_t nb = (n + kPerBlock - 1) / kPerBlock;
  int32_t *tot = tmp;              // nb + 1 entries
  k_block_totals<<<(unsigned)nb, kThreads, 0, st>>>(in, n, tot);
  PICARD_HIP(hipGetLastError());
  PICARD_TRY(scan_level(st, tot, tot, nb, tmp + nb + 1));
  k_block_scan<<<(unsigned)nb, kThreads, 0, st>>>(in, n, out, tot);
  PICARD_HIP(hipGetLastError());
  return PICARD_OK;
}

int scan_exclusive_stream(hipStream_t st, const int32_t *in, int32_t *out, int64_t n,
                          int32_t *tmp, int64_t tmp_len) {
  if (n <= 0 || scan_tmp_required(n) > tmp_len) {
    set_error("scan: empty input or temporary arena too small");
    return PICARD_ERR_STATE;
  }
  return scan_level(st, in, out, n, tmp);
}

int scan_exclusive_i32(picard_ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  if (n <= 0) {
    PICARD_HIP(hipMemsetAsync(out, 0, sizeof(int32_t), c->stream));
    return PICARD_OK;
  }
  if (scan_tmp_required(n) > c->s.scan_tmp_len) {
    set_error("scan: temporary arena too small");
    return PICARD_ERR_STATE;
  }
  return scan_level(c->stream, in, out, n, c->s.scan_tmp);
}

}  // namespace picard
