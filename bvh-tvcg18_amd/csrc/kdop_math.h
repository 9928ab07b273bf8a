// kdop_math.h -- float vector math with the reference's exact operation order.
//
// Every expression below rounds after each operation in the order the
// reference's tiny/geometry templates evaluate it, and the library is compiled
// with -ffp-contract=off, so no a*b+c is fused.  Correctly rounded sqrt and
// division come from -fhip-fp32-correctly-rounded-divide-sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>

namespace kdop_amd {

struct f3 {
  float x, y, z;
};

__host__ __device__ inline f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
// tiny_vector_ops_policy.h:33-45: ((0 + a0 b0) + a1 b1) + a2 b2
__host__ __device__ inline float dot3(const f3& a, const f3& b) {
  float d = 0.0f;
  d = d + a.x * b.x;
  d = d + a.y * b.y;
  d = d + a.z * b.z;
  return d;
}
// tiny_vector_ops_policy.h:47-54
__host__ __device__ inline f3 cross3(const f3& a, const f3& b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline f3 sub3(const f3& a, const f3& b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 add3(const f3& a, const f3& b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 mul3(const f3& a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 neg3(const f3& a) { return f3{-a.x, -a.y, -a.z}; }
// tiny_vector_ops_policy.h:57-75: v * (1 / sqrt(v.v)) when the length is > 0
__host__ __device__ inline f3 unit3(const f3& v) {
  const float l = sqrtf(dot3(v, v));
  if (l > 0.0f) {
    const float inv = 1.0f / l;
    return mul3(v, inv);
  }
  return v;
}
// std::min / std::max as the reference calls them
__host__ __device__ inline float smin(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ inline float smax(float a, float b) { return (a < b) ? b : a; }

// tiny_quaternion_functions.h:137-152 + tiny_quaternion_ops_policy.h:29-38:
// rotate(q, r) = Im(prod(prod(q, r), conj(q))), q = (x, y, z, w).
__host__ __device__ inline f3 rotate3(float qx, float qy, float qz, float qw, const f3& r) {
  const f3 u = f3{qx, qy, qz};
  const float a = -dot3(u, r);
  const f3 A = add3(cross3(u, r), mul3(r, qw));
  const f3 B = neg3(u);
  f3 o;
  o.x = (a * B.x + A.x * qw) + (A.y * B.z - A.z * B.y);
  o.y = (a * B.y + A.y * qw) + (A.z * B.x - A.x * B.z);
  o.z = (a * B.z + A.z * qw) + (A.x * B.y - A.y * B.x);
  return o;
}

}  // namespace kdop_amd
