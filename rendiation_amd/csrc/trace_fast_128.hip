// trace_fast_128.hip -- fast-traversal kernels with 128-thread blocks (one
// translation unit per block size, so the variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_fast_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<128, 0>(p, stack, t, st);
}

}  // namespace rdrt
