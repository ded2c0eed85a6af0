// pt_device.h -- device helpers of the wavefront path tracer: sample streams,
// Lambertian sampling, camera rays and hit-point offsetting.  Included by HIP
// sources only.  Integer sample generation is bit-exact with the CPU oracle.
#ifndef RDRT_PT_DEVICE_H
#define RDRT_PT_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdrt {
namespace pt {

// xxhash32 of three words (scene/rendering/gpu-ray-tracing/src/sampler.rs:33-42).
__device__ __forceinline__ uint32_t xxhash32(uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t p1 = 2246822519u, p2 = 3266489917u, p3 = 668265263u, p4 = 374761393u;
  uint32_t h = z + p4 + x * p2;
  h = p3 * ((h << 17) | (h >> 15));
  h = h + y * p2;
  h = p3 * ((h << 17) | (h >> 15));
  h = p1 * (h ^ (h >> 15));
  h = p2 * (h ^ (h >> 13));
  return h ^ (h >> 16);
}

// 24-bit fixed point to [0, 1 - 2^-23] (sobol.rs:47-50,64-67)
__device__ __forceinline__ float sample_to_float(uint32_t x) {
  return fminf((float)((x >> 8) & 0xffffffu) / 16777216.0f, 1.0f - 1.1920928955078125e-7f);
}
// scrambled base-2 radical inverse, Sobol dimension 0 (sobol.rs:40-51)
__device__ __forceinline__ float van_der_corput(uint32_t n, uint32_t scramble) {
  return sample_to_float(__builtin_bitreverse32(n) ^ scramble);
}
// scrambled Sobol dimension 1 (sobol.rs:55-68): XOR of the generator columns
// v_1 = 0x80000000, v_{k+1} = v_k ^ (v_k >> 1) for each set bit of n.
__device__ __forceinline__ float sobol1(uint32_t n, uint32_t scramble) {
  uint32_t v = 1u << 31;
  for (; n; n >>= 1, v ^= v >> 1)
    if (n & 1u) scramble ^= v;
  return sample_to_float(scramble);
}

// Per-pixel scramble of sample dimension pair k: xxhash32(px, py, seed * 2^16 + k).
__device__ __forceinline__ uint32_t scramble(uint32_t px, uint32_t py, uint32_t seed, uint32_t k) {
  return xxhash32(px, py, seed * 65536u + k);
}

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize3(V3 a) {
  float m = dot3(a, a);
  if (m > 0.0f) {
    float inv = 1.0f / sqrtf(m);
    return a * inv;
  }
  return a;
}

// sin and cos on [-pi/4, pi/4] from their Taylor series through x^9 / x^10
// (truncation below 2e-9 there, a fraction of an FP32 ulp), evaluated by
// Horner in FP32.  Written out so the oracle can run the same operations and
// sampled directions agree bit for bit.
__device__ __forceinline__ void sincos_quarter(float x, float &s, float &c) {
  const float x2 = x * x;
  s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f +
                                                                                 x2 * (-1.0f / 3628800.0f)))));
}

// concentric_sample_disk (math/statistics/src/distribution_map.rs:10-28,
// shader/library/src/sampling.rs:91-119).  In the second wedge the angle is
// pi/2 - phi with |phi| <= pi/4, so cos and sin swap roles.
__device__ __forceinline__ void concentric_disk(float s0, float s1, float &x, float &y) {
  const float PI_OVER_4 = 0.785398163397448309616f;
  float ox = s0 * 2.0f - 1.0f, oy = s1 * 2.0f - 1.0f;
  if (ox == 0.0f && oy == 0.0f) {
    x = y = 0.0f;
    return;
  }
  float sn, cs;
  if (fabsf(ox) > fabsf(oy)) {
    sincos_quarter(PI_OVER_4 * (oy / ox), sn, cs);
    x = cs * ox;
    y = sn * ox;
  } else {
    sincos_quarter(PI_OVER_4 * (ox / oy), sn, cs);
    x = sn * oy;
    y = cs * oy;
  }
}

// Orthonormal basis around n (Duff et al.; shader/library/src/sampling.rs:72-86),
// applied to a local direction.
__device__ __forceinline__ V3 tbn_apply(V3 n, V3 l) {
  float sign = n.z < 0.0f ? -1.0f : 1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  V3 t = normalize3(v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x));
  V3 bt = normalize3(v3(b, sign + n.y * n.y * a, -n.y));
  return t * l.x + bt * l.y + n * l.z;
}

// Lambertian importance sample around n (content/lighting/transport/src/
// surface/microfacet/device/diffuse.rs:24-36); z is clamped at 0.
__device__ __forceinline__ V3 sample_cosine(V3 n, float s0, float s1) {
  float x, y;
  concentric_disk(s0, s1, x, y);
  float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
  return tbn_apply(n, v3(x, y, z));
}

// offset_ray_hit (scene/rendering/gpu-ray-tracing/src/ray_util.rs:6-40):
// move the hit point off the surface by integer ulps along the normal.
__device__ __forceinline__ float offset_axis(float p, float n) {
  const float int_scale = 256.0f, float_scale = 1.0f / 65536.0f, origin = 1.0f / 32.0f;
  int of_i = (int)(n * int_scale);
  int pi = __float_as_int(p) + (p < 0.0f ? -of_i : of_i);
  return fabsf(p) < origin ? p + float_scale * n : __int_as_float(pi);
}
__device__ __forceinline__ V3 offset_ray_hit(V3 p, V3 n) {
  return v3(offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z));
}

// Camera ray (scene/rendering/gpu-ray-tracing/src/camera.rs:66-99 with
// shader/library/src/lib.rs:18-28): uv -> NDC (y up) at depth 1 through the
// inverse view-projection, direction from the camera position.
__device__ __forceinline__ V3 unproject(const float *m, float u, float v) {
  float x = u * 2.0f - 1.0f, y = (v * 2.0f - 1.0f) * -1.0f, z = 1.0f, w = 1.0f;
  float rx = x * m[0] + y * m[4] + z * m[8] + w * m[12];
  float ry = x * m[1] + y * m[5] + z * m[9] + w * m[13];
  float rz = x * m[2] + y * m[6] + z * m[10] + w * m[14];
  float rw = x * m[3] + y * m[7] + z * m[11] + w * m[15];
  return v3(rx / rw, ry / rw, rz / rw);
}

}  // namespace pt
}  // namespace rdrt

#endif
