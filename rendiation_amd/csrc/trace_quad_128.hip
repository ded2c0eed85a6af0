// trace_quad_128.hip -- fast-traversal kernels over the child-quad layout with
// 128-thread blocks (one translation unit per block size and layout, so the
// variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_quad_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<128, 1>(p, stack, t, st);
}

}  // namespace rdrt
