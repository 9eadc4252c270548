// noalloc.hip -- picard_particles_read_noalloc: the particle download of
// picard_particles_read staged through a fixed device array instead of a
// staging buffer the library allocates, for callers that cannot allocate
// device memory.  No counterpart in the reference.
//
// The array lives for the whole process on every GPU the library runs on,
// one copy per GPU.  Its device address is looked up once per GPU (with that
// GPU current) and kept; host code never passes the array's name itself.
// Every call on a GPU shares that GPU's copy, which keeps the last call's
// contents: calls that overlap on one GPU race.
#include <mutex>

#include "picard_internal.h"

namespace picard {
namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;   // 16 MiB per GPU
constexpr int64_t kBytesPerParticle = 8 * sizeof(double) + sizeof(int32_t);
constexpr int64_t kMaxRead = (int64_t)kScratchBytes / kBytesPerParticle;
static_assert(kMaxRead == PICARD_READ_NOALLOC_MAX, "header limit");
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};   // per GPU: the device address of `scratch` there

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The calling thread's current device must be `device`.
int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("particles_read_noalloc: device ordinal beyond " + std::to_string(kMaxDevices));
    return PICARD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    PICARD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("particles_read_noalloc: no device address for the staging array");
      return PICARD_ERR_HIP;
    }
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return PICARD_OK;
}

}  // namespace
}  // namespace picard

using namespace picard;

extern "C" int picard_particles_read_noalloc(picard_ctx *c, int64_t first, int64_t count,
                                             double *rp, int32_t *sp) {
  if (!c) {
    set_error("particles_read_noalloc: NULL context");
    return PICARD_ERR_ARG;
  }
  PICARD_HIP(hipSetDevice(c->device));
  if (count < 0 || (count > 0 && !rp)) {
    set_error("particles_read_noalloc: bad arguments");
    return PICARD_ERR_ARG;
  }
  if (count > kMaxRead) {
    set_error("particles_read_noalloc: at most " + std::to_string(kMaxRead) +
              " particles per call (16 MiB staging array), asked for " + std::to_string(count));
    return PICARD_ERR_ARG;
  }
  if (first < 0 || first + count > c->n) {
    set_error("particles_read_noalloc: range outside [0, num_local)");
    return PICARD_ERR_ARG;
  }
  void *base = nullptr;
  PICARD_TRY(scratch_address(c->device, &base));
  if (count == 0) return PICARD_OK;
  double *stage = static_cast<double *>(base);
  int32_t *stage_sp = reinterpret_cast<int32_t *>(stage + 8 * count);
  PICARD_TRY(stage_particles(c, first, count, stage, stage_sp));
  PICARD_HIP(hipMemcpyAsync(rp, stage, sizeof(double) * 8 * count, hipMemcpyDeviceToHost,
                            c->stream));
  if (sp)
    PICARD_HIP(hipMemcpyAsync(sp, stage_sp, sizeof(int32_t) * count, hipMemcpyDeviceToHost,
                              c->stream));
  PICARD_HIP(hipStreamSynchronize(c->stream));
  return PICARD_OK;
}
