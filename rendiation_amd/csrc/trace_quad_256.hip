// trace_quad_256.hip -- fast-traversal kernels over the child-quad layout with
// 256-thread blocks (one translation unit per block size and layout, so the
// variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_quad_256(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<256, 1>(p, stack, t, st);
}

}  // namespace rdrt
