// det_sum.hip -- kernels of the bit-reproducible accumulation (det_sum.h).
#include "det_sum.h"

#pragma clang fp contract(off)

namespace rdrt {
namespace det {
namespace {

__global__ void k_det_clear(u64 *acc, uint32_t *flags, size_t n, Header *h) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    acc[i] = 0ull;
    flags[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    h->maxbits = 0u;
    h->e = 0;
  }
}

// |x| < 2^ub for every finite term; E = 62 - ceil(log2 N) - ub keeps every
// |q| <= 2^(62 - ceil(log2 N)), so N of them sum to at most 2^62.
__global__ void k_det_scale(Header *h, uint32_t log2_terms) {
  const uint32_t m = h->maxbits;
  if (m == 0u) {
    h->e = 0;
    return;
  }
  const int32_t ef = (int32_t)(m >> 23);
  const int32_t ub = ef == 0 ? -126 : ef - 126;
  h->e = 62 - (int32_t)log2_terms - ub;
}

__global__ void k_det_finish(const u64 *acc, const uint32_t *flags, size_t n, const Header *h, float *out) {
  const double inv = pow2(-h->e);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t f = flags[i];
    float v;
    if (f & kNaN || (f & kPosInf && f & kNegInf)) v = __uint_as_float(0x7fc00000u);
    else if (f & kPosInf) v = __uint_as_float(0x7f800000u);
    else if (f & kNegInf) v = __uint_as_float(0xff800000u);
    else v = (float)((double)(long long)acc[i] * inv);
    out[i] = v;
  }
}

uint32_t grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (uint32_t)g;
}

}  // namespace

hipError_t enqueue_clear(u64 *acc, uint32_t *flags, size_t n, Header *h, hipStream_t st) {
  hipLaunchKernelGGL(k_det_clear, dim3(grid_for(n)), dim3(256), 0, st, acc, flags, n, h);
  return hipGetLastError();
}

hipError_t enqueue_scale(Header *h, uint64_t max_terms_per_entry, hipStream_t st) {
  const uint32_t l = ceil_log2(max_terms_per_entry < 1 ? 1 : max_terms_per_entry);
  hipLaunchKernelGGL(k_det_scale, dim3(1), dim3(1), 0, st, h, l);
  return hipGetLastError();
}

hipError_t enqueue_finish(const u64 *acc, const uint32_t *flags, size_t n, const Header *h, float *out,
                          hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_det_finish, dim3(grid_for(n)), dim3(256), 0, st, acc, flags, n, h, out);
  return hipGetLastError();
}

}  // namespace det
}  // namespace rdrt
