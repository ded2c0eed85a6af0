// trace_fast_256.hip -- fast-traversal kernels with 256-thread blocks (one
// translation unit per block size, so the variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_fast_256(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<256, 0>(p, stack, t, st);
}

}  // namespace rdrt
