// pt_tail_1.hip -- the path tracer's tail kernels over node layout 1
// (child quads); one translation unit per layout, so the
// variants compile in parallel.
#include "pt_tail.h"

namespace rdrt {

hipError_t launch_pt_tail_1(const PtLaunch &P, const TraceParams &pe, const TraceParams &ps, uint32_t stack,
                            const TraceTune &t, hipStream_t st) {
  return launch_tail_layout<1>(P, pe, ps, stack, t, st);
}

}  // namespace rdrt
