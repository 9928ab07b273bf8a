// kdop_single.h -- kdop::single_traversal of spheres against one body
// (SIMULATION/KDOP/KDOP/include/kdop_single_traversal.h:41-161) with the
// sphere-tetrahedron contact of contacts/geometry_contacts_sphere_tetrahedron.h.
//
// All spheres are traversed together, breadth first, over (sphere, node)
// items kept in the reference's visiting order, so each sphere's contacts come
// out in the order the sequential code reports them and the output is grouped
// by sphere without a sort.  At and below the branch-chunk roots every node
// gets the reference's overlap test and a leaf the reference's contact
// routine.  The reference tests all C branch roots per sphere; here the super
// levels above them are tested first with the same overlap test, which can
// only fail for a branch root when it fails for every super node above it (a
// super DOP contains the DOPs below).
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

struct Plane3 {
  f3 n;
  float w;
};
// types/geometry_plane.h:103-111 (unit normal of the triangle, offset p0 . n)
__device__ inline Plane3 plane3(const f3& p0, const f3& p1, const f3& p2) {
  const f3 n = unit3(cross3(sub3(p1, p0), sub3(p2, p0)));
  return Plane3{n, dot3(p0, n)};
}
// types/geometry_plane.h:156-163 (a normal and a point on the plane)
__device__ inline Plane3 plane3_np(const f3& n, const f3& p) { return Plane3{n, dot3(p, n)}; }
// types/geometry_plane.h:172-175
__device__ inline float signed_distance(const f3& p, const Plane3& P) { return dot3(P.n, p) - P.w; }
// closest_points/geometry_closest_point_on_plane.h:14-28
__device__ inline f3 closest_on_plane(const f3& p, const Plane3& P) {
  return sub3(p, mul3(P.n, dot3(P.n, sub3(p, mul3(P.n, P.w)))));
}
// closest_points/geometry_closest_point_on_line.h:16-31 with make_line(a, b)
// (types/geometry_line.h:55-70: direction unit(b - a))
__device__ inline f3 closest_on_line(const f3& p, const f3& a, const f3& b) {
  const f3 d = unit3(sub3(b, a));
  return add3(a, mul3(d, dot3(d, sub3(p, a))));
}
// geometry_inside.h:135-161 with test_face_plane = false
__device__ inline bool inside_triangle_walls(const f3& p, const f3* t) {
  const Plane3 pl = plane3(t[0], t[1], t[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const f3& p0 = t[k];
    const f3& p1 = t[(k + 1) % 3];
    const f3 p2 = add3(p1, pl.n);
    if (signed_distance(p, plane3(p0, p1, p2)) > 0.0f) return false;
  }
  return true;
}

// contacts/geometry_contacts_sphere_tetrahedron.h:35-195: at most one contact
// (true, with its point, normal and depth); surf bit v: the face opposite
// vertex v is on the surface.
__device__ inline bool sphere_tet_contact(const f3& c, float r, const f3* B, uint32_t surf, bool flip, f3& p, f3& n, float& depth) {
  // Every loop is unrolled so that each array index is a constant and the
  // arrays stay in registers (a dynamically indexed array goes to scratch
  // memory, which for tens of millions of lanes is gigabytes of traffic).
  constexpr int FACE[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};  // get_opposite_face
  Plane3 plane[4];
  float d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {  // :51-60, stopping at the first face too far away
    plane[k] = plane3(B[FACE[k][0]], B[FACE[k][1]], B[FACE[k][2]]);
    d[k] = signed_distance(c, plane[k]);
    if (d[k] > r) return false;
  }
  if (d[0] < 0.0f && d[1] < 0.0f && d[2] < 0.0f && d[3] < 0.0f) {  // :63-84 center inside
    // the first face whose distance is below no other's, taken by selects
    // (returning from inside the loop let the compiler index plane[] at run time)
    bool found = false;
    Plane3 P{f3{0.0f, 0.0f, 0.0f}, 0.0f};
    float dk = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool pick = !found && !(d[k] < d[(k + 1) % 4] || d[k] < d[(k + 2) % 4] || d[k] < d[(k + 3) % 4]);
      P.n.x = pick ? plane[k].n.x : P.n.x;
      P.n.y = pick ? plane[k].n.y : P.n.y;
      P.n.z = pick ? plane[k].n.z : P.n.z;
      P.w = pick ? plane[k].w : P.w;
      dk = pick ? d[k] : dk;
      found = found || pick;
    }
    if (found) {
      n = flip ? P.n : f3{-P.n.x, -P.n.y, -P.n.z};
      p = closest_on_plane(c, P);
      depth = dk;
      return true;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // :87-112 vertex regions
    if (!(((surf >> ((i + 1) & 3)) | (surf >> ((i + 2) & 3)) | (surf >> ((i + 3) & 3))) & 1u)) continue;
    const Plane3 v1 = plane3_np(unit3(sub3(B[i], B[(i + 1) & 3])), B[i]);
    const Plane3 v2 = plane3_np(unit3(sub3(B[i], B[(i + 2) & 3])), B[i]);
    const Plane3 v3 = plane3_np(unit3(sub3(B[i], B[(i + 3) & 3])), B[i]);
    if (signed_distance(c, v1) >= 0.0f && signed_distance(c, v2) >= 0.0f && signed_distance(c, v3) >= 0.0f) {
      const f3 m = sub3(c, B[i]);
      const float dist = sqrtf(dot3(m, m));
      if (dist > r) return false;
      const f3 u = unit3(m);
      n = flip ? u : f3{-u.x, -u.y, -u.z};
      p = B[i];
      depth = dist - r;
      return true;
    }
  }
  constexpr int EDGE[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};
#pragma unroll
  for (int v = 0; v < 6; ++v) {  // :115-160 edge regions
    const int i = EDGE[v][0], j = EDGE[v][1], k = EDGE[v][2], m = EDGE[v][3];
    if (!(((surf >> k) | (surf >> m)) & 1u)) continue;
    const f3 e = sub3(B[j], B[i]);
    const f3 nk = unit3(cross3(e, plane[k].n));
    const f3 nm = unit3(cross3(f3{-e.x, -e.y, -e.z}, plane[m].n));
    if (signed_distance(c, plane3_np(nk, B[i])) >= 0.0f && signed_distance(c, plane3_np(nm, B[i])) >= 0.0f) {
      const f3 q = closest_on_line(c, B[i], B[j]);
      const f3 mm = sub3(c, q);
      const float dist = sqrtf(dot3(mm, mm));
      if (dist > r) return false;
      const f3 u = unit3(mm);
      n = flip ? u : f3{-u.x, -u.y, -u.z};
      p = q;
      depth = dist - r;
      return true;
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {  // :163-185 face regions
    if (!((surf >> v) & 1u)) continue;
    const float dist = signed_distance(c, plane[v]);
    if (dist < 0.0f) continue;
    const f3 tri[3] = {B[FACE[v][0]], B[FACE[v][1]], B[FACE[v][2]]};
    if (inside_triangle_walls(c, tri)) {
      n = flip ? plane[v].n : f3{-plane[v].n.x, -plane[v].n.y, -plane[v].n.z};
      p = closest_on_plane(c, plane[v]);
      depth = dist - r;
      return true;
    }
  }
  return false;
}

// ----------------------------------------------------------- traversal
// Breadth first over (sphere, node) items of every sphere at once, in an
// order that is the reference's: the frontier holds, sphere by sphere, the
// untested nodes and the tested leaves left to right, and every step replaces
// each item in place by what it becomes (nothing when its test fails, the
// tested leaf itself, or its two children left child first), with the places
// taken from an exclusive scan of those counts.  Within one chunk tree the
// depth-first walk of kdop_single_traversal.h:41-117 reaches leaves left to
// right (it descends into the first child first), and the chunks are walked in
// order (:139-160), so the leaves come out in the order the reference tests
// them, which is increasing tet order.  Each node is tested once, with the
// reference's overlap test, exactly when its parent passed, as in the
// depth-first walk (a single traversal carries no state from one leaf to the
// next).  Every level of chunk trees adds at most log2 L steps below the
// top-level chunk roots (the leaves of a super chunk tree are the roots of
// the chunks below), so H log2 L steps leave only leaves.

constexpr uint32_t kSgTested = 0x80000000u;  // item node bit: a leaf whose test passed

// A few device words read by the host without a copy and a stream
// synchronisation: one thread copies them into mapped, coherent host memory,
// then releases `seq` at system scope; the host waits for that number.
struct HostWords {
  uint32_t v[4];
  uint32_t seq;
};
__global__ void k_publish_words(const uint32_t* __restrict__ src, uint32_t n, HostWords* __restrict__ dst, uint32_t seq) {
  if (threadIdx.x != 0) return;
  for (uint32_t i = 0; i < n; ++i) dst->v[i] = src[i];
  __threadfence_system();
  __hip_atomic_store(&dst->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// list[off[j]] = j for every flagged j (off: the exclusive scan of the flags):
// the flagged items in order, for a pass that works on them alone.
__global__ __launch_bounds__(256) void k_gather_flagged(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off, uint32_t m,
                                                        uint32_t* __restrict__ list) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m && flag[j]) list[off[j]] = j;
}

// geometry_convert.h:13-31: the sphere's DOP, o -/+ radius per direction.
template <int K>
__device__ inline void sphere_dop(const float* __restrict__ spheres, uint32_t i, const DirTable& dt, f3& c, float& r, float* sd) {
  const float4 s = reinterpret_cast<const float4*>(spheres)[i];
  c = f3{s.x, s.y, s.z};
  r = s.w;
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const float o = dot3(f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]}, c);
    sd[2 * k] = o - r;
    sd[2 * k + 1] = o + r;
  }
}

// Seed: the root test per sphere (kdop_single_traversal.h:139-141), then the
// top-level chunk roots of the spheres that pass.
template <int K>
__global__ __launch_bounds__(256) void k_sg_seed_count(const BodyDev* __restrict__ bodies, uint32_t body, const float* __restrict__ nodes,
                                                       DirTable dt, uint32_t n, const float* __restrict__ spheres, uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BodyDev& B = bodies[body];
  f3 c;
  float r, sd[K];
  sphere_dop<K>(spheres, i, dt, c, r, sd);
  NodeRec<K> rec;
  load_rec<K>(nodes, B.node_base, rec);
  count[i] = dop_overlap<K>(rec.v, sd) ? B.ntop : 0u;
}
__global__ __launch_bounds__(256) void k_sg_seed_write(const BodyDev* __restrict__ bodies, uint32_t body, const uint32_t* __restrict__ tops,
                                                       uint32_t n, const uint32_t* __restrict__ count, const uint32_t* __restrict__ off,
                                                       uint2* __restrict__ items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || count[i] == 0u) return;
  const BodyDev& B = bodies[body];
  const uint32_t o = off[i];
  for (uint32_t t = 0; t < B.ntop; ++t) items[o + t] = make_uint2(i, tops[B.top_base + t]);
}

// One step takes two levels of the walk: untested item j becomes nothing (its
// test fails), the tested leaf itself, or, child by child in order (first
// child first, :91-110), nothing (the child's test fails), the tested child
// leaf, or the child's two children.  So a node is tested exactly when its
// parent passed, and the frontier keeps the order described above.
// count[j] is the number of items j becomes, kind[j] which: kSelf, or 2 bits
// per child (0 nothing, 1 the tested leaf, 2 its children).  The level in
// between is never written, scanned or read back.
constexpr uint8_t kSelf = 0x10;
template <int K>
__device__ inline uint32_t child_kind(const float* __restrict__ nodes, uint32_t node, const float* sd) {
  NodeRec<K> rec;
  load_rec<K>(nodes, node, rec);
  if (!dop_overlap<K>(rec.v, sd)) return 0u;  // kdop_single_traversal.h:63-64
  return ((rec.m0 & F_BRANCH) && (rec.m0 & F_LEAF)) ? 1u : 2u;
}
template <int K>
__global__ __launch_bounds__(256) void k_sg_expand2_count(const uint2* __restrict__ items, uint32_t m, const float* __restrict__ nodes,
                                                          DirTable dt, const float* __restrict__ spheres, uint32_t* __restrict__ count,
                                                          uint8_t* __restrict__ kind) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint2 it = items[j];
  if (it.y & kSgTested) { count[j] = 1u; kind[j] = kSelf; return; }
  f3 c;
  float r, sd[K];
  sphere_dop<K>(spheres, it.x, dt, c, r, sd);
  NodeRec<K> rec;
  load_rec<K>(nodes, it.y, rec);
  if (!dop_overlap<K>(rec.v, sd)) { count[j] = 0u; kind[j] = 0u; return; }
  if ((rec.m0 & F_BRANCH) && (rec.m0 & F_LEAF)) { count[j] = 1u; kind[j] = kSelf; return; }
  const uint32_t k0 = child_kind<K>(nodes, rec.m0 & IDX_MASK, sd);  // :91-110, first child first
  const uint32_t k1 = child_kind<K>(nodes, rec.m1 & IDX_MASK, sd);
  count[j] = k0 + k1;
  kind[j] = (uint8_t)(k0 | (k1 << 2));
}
template <int K>
__global__ __launch_bounds__(256) void k_sg_expand2_write(const uint2* __restrict__ items, uint32_t m, const float* __restrict__ nodes,
                                                          const uint8_t* __restrict__ kind, const uint32_t* __restrict__ off,
                                                          uint2* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t kd = kind[j];
  if (kd == 0u) return;
  const uint2 it = items[j];
  uint32_t o = off[j];
  if (kd == kSelf) { out[o] = make_uint2(it.x, it.y | kSgTested); return; }
  const uint2 mm = load_meta<K>(nodes, it.y);
  const uint32_t ch[2] = {mm.x & IDX_MASK, mm.y & IDX_MASK};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t kh = (kd >> (2 * h)) & 3u;
    if (kh == 1u) {
      out[o++] = make_uint2(it.x, ch[h] | kSgTested);
    } else if (kh == 2u) {
      const uint2 mc = load_meta<K>(nodes, ch[h]);
      out[o++] = make_uint2(it.x, mc.x & IDX_MASK);
      out[o++] = make_uint2(it.x, mc.y & IDX_MASK);
    }
  }
}

// The leaves: the test of a leaf not yet tested, then the contact routine
// (:68-89).  write = 0: flag[j] = 1 when leaf j gives a contact (and *bad is
// set if an item is not a leaf, which the step count rules out); write = 1:
// thread k takes the flagged leaf j = list[k] (the flagged leaves gathered
// in order, so few lanes idle) and writes its contact at out[k], numbering
// spheres from sphere_base (the items number them within their batch).
template <int K>
__global__ __launch_bounds__(256) void k_sg_leaves(const BodyDev* __restrict__ bodies, uint32_t body, const uint2* __restrict__ items,
                                                   uint32_t m, const float* __restrict__ nodes, DirTable dt, const float* __restrict__ spheres,
                                                   uint32_t sphere_base, int flip, int write, uint32_t* __restrict__ flag, const uint32_t* __restrict__ list,
                                                   kdop_sphere_contact* __restrict__ out, uint32_t* __restrict__ bad) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t j = write ? list[k] : k;
  const uint2 it = items[j];
  const uint32_t node = it.y & ~kSgTested;
  f3 c;
  float r, sd[K];
  sphere_dop<K>(spheres, it.x, dt, c, r, sd);
  uint32_t m0;
  if (it.y & kSgTested) {
    m0 = load_meta<K>(nodes, node).x;
  } else {
    NodeRec<K> rec;
    load_rec<K>(nodes, node, rec);
    m0 = rec.m0;
    if (!((m0 & F_BRANCH) && (m0 & F_LEAF))) {
      if (!write) { flag[j] = 0u; atomicOr(bad, 1u); }
      return;
    }
    if (!dop_overlap<K>(rec.v, sd)) {
      if (!write) flag[j] = 0u;
      return;
    }
  }
  const BodyDev& B = bodies[body];
  const uint32_t t = m0 & IDX_MASK;
  const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)t);
  const uint32_t vi[4] = {tv.x, tv.y, tv.z, tv.w};
  f3 P[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) P[v] = world_vertex(B, vi[v]);
  f3 p, nn;
  float depth;
  const bool hit = sphere_tet_contact(c, r, P, B.surf[t], flip != 0, p, nn, depth);
  if (!write) { flag[j] = hit ? 1u : 0u; return; }
  kdop_sphere_contact& o = out[k];
  o.sphere = sphere_base + it.x;
  o.tet = t;
  o.p[0] = p.x; o.p[1] = p.y; o.p[2] = p.z;
  o.n[0] = nn.x; o.n[1] = nn.y; o.n[2] = nn.z;
  o.depth = depth;
}

}  // namespace kdop_amd
