// trace_fast_64.hip -- fast-traversal kernels with 64-thread blocks (one
// translation unit per block size, so the variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_fast_64(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<64, 0>(p, stack, t, st);
}

}  // namespace rdrt
