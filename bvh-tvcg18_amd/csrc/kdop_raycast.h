// kdop_raycast.h -- kdop::raycast (SIMULATION/KDOP/KDOP/include/kdop_raycast.h:19-155)
// for a batch of rays against one body, one thread per ray.
//
// Below (and at) the branch-chunk roots every node gets the reference's own
// test, in the reference's depth-first order, with the reference's running
// best length, so hits, lengths and ties come out as the sequential code
// gives them.  The reference walks all C branch roots of the tree per ray;
// here the super levels above them cull ranges of branch chunks first, with
// a test that can only drop chunks whose root the reference's test would
// reject or prune (see k_raycast), so the results do not change.
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

constexpr uint32_t kRayStack = 64;  // per-thread DFS stack (the host checks H log2 L + 2 <= it)

// Entry and exit parameters of the ray through the slabs of a DOP
// (raycast/geometry_raycast_dop.h:35-75; a slab the ray runs parallel to is
// skipped, VT::lowest / VT::highest are -/+FLT_MAX).
template <int K>
__device__ inline void ray_slabs(const f3& p, const f3& r, const float* vol, const DirTable& dt, float& t_min, float& t_max) {
  t_min = -FLT_MAX;
  t_max = FLT_MAX;
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const f3 d1 = f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]};
    const f3 d0 = f3{-d1.x, -d1.y, -d1.z};
    const float w0 = -vol[2 * k], w1 = vol[2 * k + 1];
    const float f0 = dot3(d0, r), f1 = dot3(d1, r);
    if (f0 == 0.0f) continue;
    float t_near = (w0 - dot3(d0, p)) / f0;
    float t_far = (w1 - dot3(d1, p)) / f1;
    if (t_near > t_far) {
      const float tmp = t_far;
      t_far = t_near;
      t_near = tmp;
    }
    t_min = (t_near > t_min) ? t_near : t_min;
    t_max = (t_far < t_max) ? t_far : t_max;
  }
}

// compute_raycast_dop with a threshold (raycast/geometry_raycast_dop.h:19-99,
// outside_dop of geometry_inside.h:19-40); length: the entry parameter.
template <int K>
__device__ inline bool raycast_dop(const f3& p, const f3& r, const float* vol, const DirTable& dt, float threshold, float& length) {
  float t_min, t_max;
  ray_slabs<K>(p, r, vol, dt, t_min, t_max);
  length = FLT_MAX;
  if (!(t_min > 0.0f && t_min <= t_max)) return false;
  const f3 q = add3(p, mul3(r, t_min));
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const f3 d = f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]};
    const float o = dot3(d, q);
    if (o < vol[2 * k] - threshold) return false;
    if (o > vol[2 * k + 1] + threshold) return false;
  }
  length = t_min;
  return true;
}

// compute_raycast_tetrahedron (raycast/geometry_raycast_tetrahedron.h:27-131):
// per face (get_opposite_face, types/geometry_tetrahedron.h:231-257) the
// largest of the three edge-cross normals, then the front-facing surface
// faces the ray crosses inside the other three planes; the nearest wins.
__device__ inline bool raycast_tet(const f3& o, const f3& r, const f3* P, uint32_t surf, f3& hit, float& length) {
  constexpr uint32_t FACE = (1u << 0) | (2u << 2) | (3u << 4)      // face 0: 1 2 3
                          | (0u << 6) | (3u << 8) | (2u << 10)     // face 1: 0 3 2
                          | (0u << 12) | (1u << 14) | (3u << 16)   // face 2: 0 1 3
                          | (0u << 18) | (2u << 20) | (1u << 22);  // face 3: 0 2 1
  length = FLT_MAX;
  hit = f3{0.0f, 0.0f, 0.0f};
  bool status = false;
  f3 n[4];
  float w[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const f3 a = P[(FACE >> (6 * m)) & 3u], b = P[(FACE >> (6 * m + 2)) & 3u], c = P[(FACE >> (6 * m + 4)) & 3u];
    const f3 n0 = cross3(sub3(b, a), sub3(c, a));
    const f3 n1 = cross3(sub3(c, b), sub3(a, b));
    const f3 n2 = cross3(sub3(a, c), sub3(b, c));
    const float w0 = dot3(n0, a), w1 = dot3(n1, b), w2 = dot3(n2, c);
    const float t0 = dot3(n0, n0), t1 = dot3(n1, n1), t2 = dot3(n2, n2);
    n[m] = f3{0.0f, 0.0f, 0.0f};
    w[m] = 0.0f;
    if (t0 >= t1 && t0 >= t2) { n[m] = n0; w[m] = w0; }
    else if (t1 >= t0 && t1 >= t2) { n[m] = n1; w[m] = w1; }
    else if (t2 >= t0 && t2 >= t1) { n[m] = n2; w[m] = w2; }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    if (!((surf >> m) & 1u)) continue;
    const float rn = dot3(r, n[m]);
    const float on = dot3(o, n[m]);
    if (rn >= 0.0f) continue;
    const float t = (w[m] - on) / rn;
    if (t < 0.0f) continue;
    const f3 q = add3(o, mul3(r, t));
    const int i = (m + 1) & 3, j = (m + 2) & 3, k = (m + 3) & 3;
    if (dot3(n[i], q) - w[i] > 0.0f) continue;
    if (dot3(n[j], q) - w[j] > 0.0f) continue;
    if (dot3(n[k], q) - w[k] > 0.0f) continue;
    if (t < length) {
      length = t;
      hit = q;
      status = true;
    }
  }
  return status;
}

// One thread per ray (make_ray normalises the direction, types/geometry_ray.h:39-42).
// The root test is the reference's (kdop_raycast.h:126-127).  Then the tree is
// walked depth first from the top-level chunk roots in order; a super node's
// children are the lower chunks' roots in order, so branch chunks come in
// increasing order, as kdop_raycast.h:129-149 visits them.  A super node is
// skipped when the ray's slab interval through it is empty (t_min > t_max),
// behind the origin (t_max <= 0) or starts at or beyond the best length so
// far: a DOP inside it gives, slab by slab, entry parameters no smaller and
// exit parameters no larger (the subtractions and divisions are correctly
// rounded, hence monotone), so each of its branch roots would fail the
// reference's test or be pruned by `t >= length` (:55-57) without changing
// the hit.  At branch nodes the reference's test and recursion run unchanged.
template <int K>
__global__ __launch_bounds__(64) void k_raycast(const BodyDev* __restrict__ bodies, uint32_t body, const uint32_t* __restrict__ tops,
                                                const float* __restrict__ nodes, DirTable dt, uint32_t n, const float* __restrict__ orig,
                                                const float* __restrict__ dir, float* __restrict__ hits, float* __restrict__ lengths,
                                                uint8_t* __restrict__ flags) {
  __shared__ uint32_t stk[64][kRayStack];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* st = stk[threadIdx.x];
  const BodyDev& B = bodies[body];
  const f3 o =f3{orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
  const f3 r = unit3(f3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]});
  const float threshold = 0.01f;
  f3 hit = f3{0.0f, 0.0f, 0.0f};
  float length = FLT_MAX;
  float t;
  NodeRec<K> rec;
  load_rec<K>(nodes, B.node_base, rec);
  if (raycast_dop<K>(o, r, rec.v, dt, threshold, t)) {
    for (uint32_t top = 0; top < B.ntop; ++top) {
      uint32_t sp = 0;
      st[sp++] = tops[B.top_base + top];
      while (sp) {
        const uint32_t s = st[--sp];
        load_rec<K>(nodes, s, rec);
        const uint32_t c0 = rec.m0 & IDX_MASK, c1 = rec.m1 & IDX_MASK;
        if (!(rec.m0 & F_BRANCH)) {  // super node: cull only
          float t_min, t_max;
          ray_slabs<K>(o, r, rec.v, dt, t_min, t_max);
          if (t_min > t_max || t_max <= 0.0f || t_min >= length) continue;
          st[sp++] = c1;
          st[sp++] = c0;
          continue;
        }
        if (!raycast_dop<K>(o, r, rec.v, dt, threshold, t)) continue;  // kdop_raycast.h:52-57
        if (t >= length) continue;
        if (rec.m0 & F_LEAF) {  // :61-79
          const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)c0);
          const uint32_t vi[4] = {tv.x, tv.y, tv.z, tv.w};
          f3 P[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) P[v] = f3{B.X[vi[v]], B.Y[vi[v]], B.Z[vi[v]]};
          f3 q;
          float len;
          if (raycast_tet(o, r, P, B.surf[c0], q, len)) {
            hit = len < length ? q : hit;
            length = len < length ? len : length;
          }
        } else {
          st[sp++] = c1;
          st[sp++] = c0;
        }
      }
    }
  }
  hits[3 * i] = hit.x;
  hits[3 * i + 1] = hit.y;
  hits[3 * i + 2] = hit.z;
  lengths[i] = length;
  flags[i] = length < FLT_MAX ? 1u : 0u;
}

}  // namespace kdop_amd
