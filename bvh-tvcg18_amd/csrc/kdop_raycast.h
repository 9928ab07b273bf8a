// kdop_raycast.h -- kdop::raycast (SIMULATION/KDOP/KDOP/include/kdop_raycast.h:19-155)
// for a batch of rays against one body.
//
// Below (and at) the branch-chunk roots every node gets the reference's own
// test, in the reference's depth-first order, with the reference's running
// best length, so hits, lengths and ties come out as the sequential code
// gives them.  The reference walks all C branch roots of the tree per ray;
// here the super levels above them cull chunks first, with a test that can
// only drop chunks whose root the reference's test would reject, and the
// chunks are walked in parallel before each ray's running length is applied
// (see "rays" below for why that keeps the result).
#pragma once

#include "kdop_kernels.h"
#include "kdop_single.h"

namespace kdop_amd {

// Entry and exit parameters of the ray through the slabs of a DOP
// (raycast/geometry_raycast_dop.h:35-75; a slab the ray runs parallel to is
// skipped, VT::lowest / VT::highest are -/+FLT_MAX).
//
// The reference divides (w - d.p) / (d.r) in float at every node.  The ray's
// part of each axis is fixed for a whole walk, so it is computed once per ray
// (RaySlabs): d0.p and d1.p as the reference computes them, and y = RN(1/f1),
// the double reciprocal of f1 = d1.r (0 marks a slab the ray runs parallel
// to; 1/f1 of a nonzero float is never 0 in double).  d0 = -d1 makes f0 = -f1
// bit for bit (negation commutes with every rounded product and sum), so
// 1/f0 = -y exactly.  Each quotient is then float(double(num) * y):
//  - double(num) * y is within 2^-52 (relative) of num / f;
//  - a quotient of two 24-bit floats is never a float rounding midpoint and
//    lies at least 2^-49 (relative) from one: with num = A 2^i, f = B 2^j and
//    a midpoint M 2^k (A, B < 2^24, M < 2^25 odd), A 2^i - B M 2^k' is a
//    nonzero multiple of 2^k', so the distance is at least 1 / (B M);
// so the double product rounds to the same float as the IEEE division,
// including overflow to infinity and subnormal quotients.
template <int K>
struct RaySlabs {
  float dp0[K / 2], dp1[K / 2];
  double y[K / 2];
};
template <int K>
__device__ inline void ray_slab_setup(const f3& p, const f3& r, const DirTable& dt, RaySlabs<K>& S) {
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const f3 d1 = f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]};
    const f3 d0 = f3{-d1.x, -d1.y, -d1.z};
    const float f1 = dot3(d1, r);
    S.dp0[k] = dot3(d0, p);
    S.dp1[k] = dot3(d1, p);
    S.y[k] = f1 == 0.0f ? 0.0 : 1.0 / (double)f1;
  }
}
template <int K>
__device__ inline void ray_slabs(const RaySlabs<K>& S, const float* vol, float& t_min, float& t_max) {
  t_min = -FLT_MAX;
  t_max = FLT_MAX;
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const double y = S.y[k];
    if (y == 0.0) continue;  // f0 == 0
    const float w0 = -vol[2 * k], w1 = vol[2 * k + 1];
    float t_near = (float)((double)(w0 - S.dp0[k]) * -y);  // (w0 - d0.p) / f0
    float t_far = (float)((double)(w1 - S.dp1[k]) * y);    // (w1 - d1.p) / f1
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
__device__ inline bool raycast_dop(const f3& p, const f3& r, const RaySlabs<K>& S, const float* vol, const DirTable& dt, float threshold,
                                   float& length) {
  float t_min, t_max;
  ray_slabs<K>(S, vol, t_min, t_max);
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
    const bool better = t < length;  // per-component selects keep q out of scratch memory
    hit.x = better ? q.x : hit.x;
    hit.y = better ? q.y : hit.y;
    hit.z = better ? q.z : hit.z;
    length = better ? t : length;
    status = status || better;
  }
  return status;
}

// ---------------------------------------------------------------- rays
// kdop::raycast (kdop_raycast.h:102-155): the root test, then every branch
// chunk in order, each walked depth first from its root (:19-100) with the
// running best length, which prunes nodes whose DOP entry is not below it.
// Here the walk is cut at the chunks without changing its result:
//   A. the branch-chunk roots each ray may reach, in chunk order, breadth
//      first over (ray, node) items of all rays; a super node is dropped when
//      the ray's slab interval through it is empty or behind the origin (a
//      DOP inside it then gives entry parameters no smaller and exit
//      parameters no larger, slab by slab, since correctly rounded
//      subtraction and division are monotone, so every branch root below it
//      fails the reference's DOP test);
//   B. every (ray, chunk) walked on its own from length FLT_MAX, giving the
//      chunk's best length h;
//   C. per ray, its chunks in order with the running length L: a chunk with
//      h >= L changes nothing, so only chunks with h < L are walked again,
//      from L, exactly as the reference walks them.
// Why C may skip a chunk with h >= L: walked from L, every node's prune test
// compares with a length that is at most L <= h, while the walk from
// FLT_MAX never went below h; so the walk from L prunes at least wherever the
// walk from FLT_MAX did, reaches a subset of its leaves, and finds no hit
// below h >= L.

constexpr uint32_t kChunkStack = 16;  // a chunk walk holds at most depth + 1 <= log2 L + 1 entries: L <= 2^15

// The reference's walk below one branch-chunk root (kdop_raycast.h:46-100)
// from `length`: returns the best length, hit updated with it.
template <int K>
__device__ inline float chunk_raycast(const float* __restrict__ nodes, const BodyDev& B, uint32_t root, const f3& o, const f3& r,
                                      const RaySlabs<K>& S, const DirTable& dt, float length, f3& hit, uint32_t* st) {
  uint32_t sp = 0;
  st[sp++] = root;
  NodeRec<K> rec;
  float t;
  while (sp) {
    const uint32_t s = st[--sp];
    load_rec<K>(nodes, s, rec);
    if (!raycast_dop<K>(o, r, S, rec.v, dt, 0.01f, t)) continue;  // :52-57
    if (t >= length) continue;
    const uint32_t c0 = rec.m0 & IDX_MASK, c1 = rec.m1 & IDX_MASK;
    if (rec.m0 & F_LEAF) {  // :61-79
      const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)c0);
      const uint32_t vi[4] = {tv.x, tv.y, tv.z, tv.w};
      f3 P[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) P[v] = world_vertex(B, vi[v]);
      f3 q;
      float len;
      if (raycast_tet(o, r, P, B.surf[c0], q, len)) {
        const bool better = len < length;
        hit.x = better ? q.x : hit.x;
        hit.y = better ? q.y : hit.y;
        hit.z = better ? q.z : hit.z;
        length = better ? len : length;
      }
    } else {  // :81-99, first child first
      st[sp++] = c1;
      st[sp++] = c0;
    }
  }
  return length;
}

__device__ inline void load_ray(const float* __restrict__ orig, const float* __restrict__ dir, uint32_t i, f3& o, f3& r) {
  o = f3{orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
  r = unit3(f3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]});  // make_ray (types/geometry_ray.h:39-42)
}

// A: the root test (kdop_raycast.h:126-127), then the top-level chunk roots.
template <int K>
__global__ __launch_bounds__(256) void k_rc_seed_count(const BodyDev* __restrict__ bodies, uint32_t body, const float* __restrict__ nodes,
                                                       DirTable dt, uint32_t n, const float* __restrict__ orig, const float* __restrict__ dir,
                                                       uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BodyDev& B = bodies[body];
  f3 o, r;
  load_ray(orig, dir, i, o, r);
  NodeRec<K> rec;
  load_rec<K>(nodes, B.node_base, rec);
  RaySlabs<K> S;
  ray_slab_setup<K>(o, r, dt, S);
  float t;
  count[i] = raycast_dop<K>(o, r, S, rec.v, dt, 0.01f, t) ? B.ntop : 0u;
}

// A, one step of two levels (as k_sg_expand2_count for the single traversal,
// with the slab test culling super nodes and branch-chunk roots as the leaves):
// a marked item stays; an unmarked branch-chunk root stays, marked; a super
// node whose slab interval is empty or behind the origin becomes nothing, and
// otherwise each child in order becomes nothing (its interval is empty or
// behind), itself marked (a branch-chunk root), or its two children.  count /
// kind as k_sg_expand2_write reads them.
template <int K>
__device__ inline uint32_t rc_child_kind(const float* __restrict__ nodes, uint32_t node, const RaySlabs<K>& S) {
  NodeRec<K> rec;
  load_rec<K>(nodes, node, rec);
  if (rec.m0 & F_BRANCH) return 1u;
  float t_min, t_max;
  ray_slabs<K>(S, rec.v, t_min, t_max);
  return (t_min > t_max || t_max <= 0.0f) ? 0u : 2u;
}
template <int K>
__global__ __launch_bounds__(256) void k_rc_expand2_count(const uint2* __restrict__ items, uint32_t m, const float* __restrict__ nodes,
                                                          DirTable dt, const float* __restrict__ orig, const float* __restrict__ dir,
                                                          uint32_t* __restrict__ count, uint8_t* __restrict__ kind) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint2 it = items[j];
  if (it.y & kSgTested) { count[j] = 1u; kind[j] = kSelf; return; }
  NodeRec<K> rec;
  load_rec<K>(nodes, it.y, rec);
  if (rec.m0 & F_BRANCH) { count[j] = 1u; kind[j] = kSelf; return; }
  f3 o, r;
  load_ray(orig, dir, it.x, o, r);
  RaySlabs<K> S;
  ray_slab_setup<K>(o, r, dt, S);
  float t_min, t_max;
  ray_slabs<K>(S, rec.v, t_min, t_max);
  if (t_min > t_max || t_max <= 0.0f) { count[j] = 0u; kind[j] = 0u; return; }
  const uint32_t k0 = rc_child_kind<K>(nodes, rec.m0 & IDX_MASK, S);
  const uint32_t k1 = rc_child_kind<K>(nodes, rec.m1 & IDX_MASK, S);
  count[j] = k0 + k1;
  kind[j] = (uint8_t)(k0 | (k1 << 2));
}

// B, first pass: the reference's test of each chunk root (from FLT_MAX);
// pass[j] = 1 when the root passes, and best[j] = FLT_MAX for every item
// (*bad set if an item is not a branch-chunk root, which the step count
// rules out).  Most roots fail, so the walks below run on the passing items
// only, gathered by a scan, instead of leaving most lanes of a wave idle.
template <int K>
__global__ __launch_bounds__(256) void k_rc_roots(const uint2* __restrict__ items, uint32_t m, const float* __restrict__ nodes, DirTable dt,
                                                  const float* __restrict__ orig, const float* __restrict__ dir, uint32_t* __restrict__ pass,
                                                  float* __restrict__ best, uint32_t* __restrict__ bad) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint2 it = items[j];
  const uint32_t root = it.y & ~kSgTested;
  best[j] = FLT_MAX;
  NodeRec<K> rec;
  load_rec<K>(nodes, root, rec);
  if (!(rec.m0 & F_BRANCH)) {
    pass[j] = 0u;
    atomicOr(bad, 1u);
    return;
  }
  f3 o, r;
  load_ray(orig, dir, it.x, o, r);
  RaySlabs<K> S;
  ray_slab_setup<K>(o, r, dt, S);
  float t;
  pass[j] = (raycast_dop<K>(o, r, S, rec.v, dt, 0.01f, t) && t < FLT_MAX) ? 1u : 0u;  // kdop_raycast.h:52-57
}
// B, second pass: best[j] = the best length of chunk j's walk from FLT_MAX,
// for the items in list.
template <int K>
__global__ __launch_bounds__(256) void k_rc_chunks(const BodyDev* __restrict__ bodies, uint32_t body, const uint2* __restrict__ items,
                                                   const uint32_t* __restrict__ list, uint32_t np, const float* __restrict__ nodes, DirTable dt,
                                                   const float* __restrict__ orig, const float* __restrict__ dir, float* __restrict__ best) {
  __shared__ uint32_t stk[256][kChunkStack];
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const uint32_t j = list[k];
  const uint2 it = items[j];
  f3 o, r, hit{0.0f, 0.0f, 0.0f};
  load_ray(orig, dir, it.x, o, r);
  RaySlabs<K> S;
  ray_slab_setup<K>(o, r, dt, S);
  best[j] = chunk_raycast<K>(nodes, bodies[body], it.y & ~kSgTested, o, r, S, dt, FLT_MAX, hit, stk[threadIdx.x]);
}

// first[ray] = its first item (items are grouped by ray; rays without items
// keep the ~0u the host filled in).
__global__ __launch_bounds__(256) void k_rc_first(const uint2* __restrict__ items, uint32_t m, uint32_t* __restrict__ first) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t ray = items[j].x;
  if (j == 0u || items[j - 1u].x != ray) first[ray] = j;
}

// C: per ray, its chunks in order from length FLT_MAX (VT::infinity), each
// with best < length walked again from length.  One wave per ray: the wave
// reads 64 of the ray's chunk bests at a time and a ballot finds the next
// chunk in order whose best is below the running length; lane 0 walks it
// and the length is re-broadcast, so chunks are taken in the sequential
// order against the length the sequential code would hold.  A miss leaves
// the point (0, 0, 0), length FLT_MAX and flag 0.  Rays are numbered from
// ray_base in the outputs (the items number them within their batch).
template <int K>
__global__ __launch_bounds__(256) void k_rc_resolve(const BodyDev* __restrict__ bodies, uint32_t body, const uint2* __restrict__ items,
                                                    uint32_t m, const float* __restrict__ best, const uint32_t* __restrict__ first, uint32_t n,
                                                    const float* __restrict__ nodes, DirTable dt, const float* __restrict__ orig,
                                                    const float* __restrict__ dir, uint32_t ray_base, float* __restrict__ hits,
                                                    float* __restrict__ lengths, uint8_t* __restrict__ flags) {
  __shared__ uint32_t stk[4][kChunkStack];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 4u + w;  // this wave's ray
  if (i >= n) return;
  f3 hit{0.0f, 0.0f, 0.0f};
  float length = FLT_MAX;
  const uint32_t j0 = first[i];
  if (j0 != ~0u) {
    const BodyDev& B = bodies[body];
    f3 o, r;
    load_ray(orig, dir, i, o, r);
    RaySlabs<K> S;
    ray_slab_setup<K>(o, r, dt, S);
    for (uint32_t base = j0;; base += 64u) {
      const uint32_t j = base + lane;
      const bool mine = j < m && items[j].x == i;
      const float bj = mine ? best[j] : FLT_MAX;
      const unsigned long long in_ray = __ballot(mine);
      unsigned long long done = 0ull;  // lanes already taken in this group
      for (;;) {
        const unsigned long long cand = __ballot(mine && bj < length) & ~done;
        if (!cand) break;
        const uint32_t k = (uint32_t)(__ffsll(cand) - 1);
        done |= (2ull << k) - 1ull;  // k and every lane before it
        if (lane == 0u) {
          const uint32_t root = items[base + k].y & ~kSgTested;
          length = chunk_raycast<K>(nodes, B, root, o, r, S, dt, length, hit, stk[w]);
        }
        length = readlane_f(length, 0);
      }
      if (in_ray != ~0ull) break;  // the ray's items end in this group
    }
  }
  if (lane == 0u) {
    const uint32_t g = ray_base + i;
    hits[3 * g] = hit.x;
    hits[3 * g + 1] = hit.y;
    hits[3 * g + 2] = hit.z;
    lengths[g] = length;
    flags[g] = length < FLT_MAX ? 1u : 0u;
  }
}

}  // namespace kdop_amd
