// det_sum.h -- bit-reproducible sums of float terms scattered by many threads.
//
// Float atomics add in whatever order the hardware serves them, so two runs
// over the same terms can round differently.  Here every term x is rounded
// once to the integer q = rint(x * 2^E) and the q are added with 64-bit
// integer atomics, which commute and associate exactly: the sum is the same
// whatever the order.  E is chosen per call from the largest finite |term|
// (an order-independent atomic max) and an upper bound N on the terms that
// fall on one entry, so that |q| <= 2^(62 - ceil(log2 N)) and no partial sum
// can leave the int64 range.  Each term is then exact to 2^-E, i.e. to
// max|term| * 2^-(61 - ceil(log2 N)) (about 1e-10 of the largest term at
// N = 2^27).  A non-finite term is kept aside per entry: a NaN, or both
// infinities, make the entry NaN, one infinity makes it that infinity, as
// float addition would.
//
// Protocol per accumulation (all on one stream):
//   k_det_clear -> terms: det_note_max (pass 1) -> k_det_scale ->
//   terms: det_add (pass 2) -> k_det_finish.
#ifndef RDRT_DET_SUM_H
#define RDRT_DET_SUM_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rdrt {
namespace det {

using u64 = unsigned long long;

// Device header of one accumulation.
struct Header {
  uint32_t maxbits;  // largest finite |term| as float bits (0: none yet)
  int32_t e;         // scale exponent E
};

constexpr uint32_t kPosInf = 1u, kNegInf = 2u, kNaN = 4u;

__device__ __forceinline__ double pow2(int32_t e) {
  return __longlong_as_double((long long)(e + 1023) << 52);  // |e| <= 1022 here
}

// |x| as float bits, 0 for a non-finite x
__device__ __forceinline__ uint32_t abs_bits(float x) {
  const uint32_t b = __float_as_uint(x) & 0x7fffffffu;
  return b < 0x7f800000u ? b : 0u;
}

// Wave-wide max of per-lane bits, then one atomic per wave.
__device__ __forceinline__ void det_note_max(Header *h, uint32_t bits) {
  for (int off = 32; off > 0; off >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, off, 64));
  if ((threadIdx.x & 63u) == 0 && bits) atomicMax(&h->maxbits, bits);
}

__device__ __forceinline__ uint32_t special_of(float x) {
  if (x != x) return kNaN;
  return x > 0.0f ? kPosInf : kNegInf;
}

// One term into entry i (acc: int64 sums as u64, flags: non-finite marks).
__device__ __forceinline__ void det_add(u64 *acc, uint32_t *flags, size_t i, float x, int32_t e) {
  if (x == 0.0f) return;
  if ((__float_as_uint(x) & 0x7f800000u) == 0x7f800000u) {
    atomicOr(flags + i, special_of(x));
    return;
  }
  const long long q = __double2ll_rn((double)x * pow2(e));
  atomicAdd(acc + i, (u64)q);
}

// The integer a term adds (callers that sum block-locally first).
__device__ __forceinline__ long long det_quant(float x, int32_t e) { return __double2ll_rn((double)x * pow2(e)); }
__device__ __forceinline__ bool det_finite(float x) { return (__float_as_uint(x) & 0x7f800000u) != 0x7f800000u; }

// ceil(log2(n)) for n >= 1
inline uint32_t ceil_log2(uint64_t n) {
  uint32_t k = 0;
  while ((1ull << k) < n) k++;
  return k;
}

// Host side (det_sum.hip): zero n sums and flags and the header (a plain
// kernel, so it can sit in a captured graph); set E from the noted max and
// the bound on terms per entry; write out[i] = sum_i * 2^-E as float, or
// the non-finite value the flags call for.
hipError_t enqueue_clear(u64 *acc, uint32_t *flags, size_t n, Header *h, hipStream_t st);
hipError_t enqueue_scale(Header *h, uint64_t max_terms_per_entry, hipStream_t st);
hipError_t enqueue_finish(const u64 *acc, const uint32_t *flags, size_t n, const Header *h, float *out,
                          hipStream_t st);

}  // namespace det
}  // namespace rdrt

#endif
