// trace_qquad_128.hip -- fast-traversal kernels over the quantized child-quad
// layout with 128-thread blocks (one translation unit per block size and
// layout, so the variants compile in parallel).
#include "trace_device.h"

namespace rdrt {

hipError_t launch_qquad_128(const TraceParams &p, uint32_t stack, const TraceTune &t, hipStream_t st) {
  return launch_fast_block<128, 2>(p, stack, t, st);
}

}  // namespace rdrt
