// api_static.hip -- rdrt_compact_static: the stable stream compaction of
// rdrt_compact with no device allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).  Its scratch (the
// block-local runs, block counts and offsets of the three-pass compaction)
// is one file-scope device array of fixed size, sized for
// RDRT_COMPACT_STATIC_MAX elements.  Each device the library runs on holds
// its own copy until the process ends; every call on one device shares it,
// so calls that overlap on one device race.  Nothing in it needs clearing:
// every word a call reads was written earlier in the same call.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "context.h"

using namespace rdrt;

namespace {

constexpr uint32_t kMaxN = RDRT_COMPACT_STATIC_MAX;
constexpr size_t kBlocks = ((size_t)kMaxN + kPtBlock - 1) / kPtBlock;
// block-local runs (kBlocks * kPtBlock) + block counts + block offsets + 1
constexpr size_t kWords = kBlocks * kPtBlock + 2 * kBlocks + 1;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
bool g_tried[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) uint32_t scratch[kWords];

namespace rdrt {

// Looks up this device's copy of `scratch` once (the device must be current).
void static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_tried[device]) return;
  g_tried[device] = true;
  void *p = nullptr;
  g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (g_err[device] == hipSuccess) g_ptr[device] = p;
}

}  // namespace rdrt

extern "C" {

int32_t rdrt_compact_static(rdrt_ctx *c, const uint32_t *mask, uint32_t n, uint32_t *out_idx, uint32_t *count) {
  if (!c || !count || (n && (!mask || !out_idx))) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  if (n > kMaxN)
    return fail(RDRT_ERR_CAPACITY, "rdrt_compact_static takes at most " + std::to_string(kMaxN) +
                                       " elements (its device scratch has a fixed size); use rdrt_compact");
  int32_t rc = need_device(c);
  if (rc) return rc;
  const int dev = c->cfg.device;
  void *p = nullptr;
  hipError_t err = hipErrorInvalidDevice;
  if (dev >= 0 && dev < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_mu);
    p = g_ptr[dev];
    err = g_err[dev];
  }
  if (!p)
    return fail(RDRT_ERR_STATE, std::string("the static compaction scratch of this device is unavailable: ") +
                                    hipGetErrorString(err));
  RDRT_HIP(enqueue_compact(mask, n, out_idx, count, (uint32_t *)p, c->stream));
  return RDRT_OK;
}

}  // extern "C"
