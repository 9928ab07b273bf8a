// kdop_narrow.h -- tet-tet narrow phase (MOST_OPPOSING_SURFACES) on the device.
//
// Restates geometry::contacts_tetrahedron_tetrahedron(..., MOST_OPPOSING_SURFACES)
// (FOUNDATION/GEOMETRY/GEOMETRY/include/contacts/geometry_contacts_tetrahedron_tetrahedron.h:1312-1345):
// SAT gate (overlap/geometry_overlap_tetrahedron_tetrahedron.h:50-290), the most
// opposing surface normal (:464-533) and contact generation from the
// intersection (:543-681), in FP32 with the reference's operation order.
#pragma once

#include "kdop_math.h"

namespace kdop_amd {

// Small index tables of the reference, packed two bits per entry so that a
// lane-dependent lookup is a shift, not a memory access.
template <int N>
constexpr unsigned long long pack2(const int (&t)[N]) {
  unsigned long long v = 0;
  for (int i = 0; i < N; ++i) v |= (unsigned long long)t[i] << (2 * i);
  return v;
}
// faces opposite vertex 0..3 (types/geometry_tetrahedron.h:231-257)
constexpr int kFaceT[12] = {1, 2, 3, 0, 3, 2, 0, 1, 3, 0, 2, 1};
// edges 01 02 03 12 13 23 (overlap/geometry_overlap_tetrahedron_tetrahedron.h:102-107)
constexpr int kEdgeT[12] = {0, 1, 0, 2, 0, 3, 1, 2, 1, 3, 2, 3};
// Vertex triples SAT_line projects for each edge-edge axis, as the reference
// hard-codes them (overlap ...tetrahedron.h:109-287): triple ids 0={1,2,3}
// 1={0,1,3} 2={0,1,2}; A side by A edge, B side by (min(A edge, 2), B edge).
constexpr int kTripleT[9] = {1, 2, 3, 0, 1, 3, 0, 1, 2};
constexpr int kSatAT[6] = {0, 0, 0, 1, 2, 2};
constexpr int kSatBT[18] = {0, 1, 2, 1, 2, 2, 0, 0, 2, 1, 2, 2, 0, 0, 0, 1, 2, 2};
constexpr unsigned long long kFaceP = pack2(kFaceT), kEdgeP = pack2(kEdgeT), kTripleP = pack2(kTripleT);
constexpr unsigned long long kSatAP = pack2(kSatAT), kSatBP = pack2(kSatBT);
__device__ inline uint32_t tab2(unsigned long long t, uint32_t i) { return (uint32_t)(t >> (2 * i)) & 3u; }
__device__ inline uint32_t face_v(uint32_t f, uint32_t k) { return tab2(kFaceP, 3 * f + k); }
__device__ inline uint32_t edge_v(uint32_t e, uint32_t k) { return tab2(kEdgeP, 2 * e + k); }
__device__ inline uint32_t triple_v(uint32_t t, uint32_t k) { return tab2(kTripleP, 3 * t + k); }
__device__ inline uint32_t sat_a(uint32_t ea) { return tab2(kSatAP, ea); }
__device__ inline uint32_t sat_b(uint32_t ea, uint32_t eb) { return tab2(kSatBP, 6 * (ea < 2u ? ea : 2u) + eb); }

struct UPlane {  // unscaled plane: normal + a point on it
  f3 n, p;
};

// contacts ...tetrahedron.h:30-62 (later ties overwrite earlier ones)
__device__ inline UPlane make_uplane(const f3& p0, const f3& p1, const f3& p2) {
  const f3 m0 = cross3(sub3(p1, p0), sub3(p2, p0));
  const f3 m1 = cross3(sub3(p2, p1), sub3(p0, p1));
  const f3 m2 = cross3(sub3(p0, p2), sub3(p1, p2));
  const float l0 = dot3(m0, m0), l1 = dot3(m1, m1), l2 = dot3(m2, m2);
  UPlane pl{f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}};
  if (l0 >= l1 && l0 >= l2) { pl.n = m0; pl.p = p0; }
  if (l1 >= l0 && l1 >= l2) { pl.n = m1; pl.p = p1; }
  if (l2 >= l1 && l2 >= l0) { pl.n = m2; pl.p = p2; }
  return pl;
}
__device__ inline float usd(const f3& q, const UPlane& pl) { return dot3(pl.n, sub3(q, pl.p)); }
__device__ inline bool inside_planes(const f3& q, const UPlane* P) {
  const float d0 = usd(q, P[0]), d1 = usd(q, P[1]), d2 = usd(q, P[2]), d3 = usd(q, P[3]);
  return !(d0 > 0.0f) && !(d1 > 0.0f) && !(d2 > 0.0f) && !(d3 > 0.0f);
}
__device__ inline bool is_crossing(const f3& a, const f3& b, const UPlane& pl) {
  const float dA = usd(a, pl), dB = usd(b, pl);
  if (dA >= 0.0f && dB >= 0.0f) return false;
  if (dA <= 0.0f && dB <= 0.0f) return false;
  return true;
}
__device__ inline f3 make_intersection(const f3& a, const f3& b, const UPlane& pl) {
  const float dA = fabsf(usd(a, pl));
  const float dB = fabsf(usd(b, pl));
  const float t = dA / (dA + dB);
  return add3(a, mul3(sub3(b, a), t));
}

// types/geometry_plane.h:103-111 with the constructor's second unit() (:66-71)
__device__ inline void make_plane(const f3& p0, const f3& p1, const f3& p2, f3& n, float& w) {
  const f3 normal = unit3(cross3(sub3(p1, p0), sub3(p2, p0)));
  w = dot3(p0, normal);
  n = unit3(normal);
}

__device__ inline bool sat_line(const f3& a0, const f3& a1, const f3& a2, const f3& b0, const f3& b1, const f3& b2, const f3& n) {
  const float t0_d0 = dot3(n, a0), t0_d1 = dot3(n, a1), t0_d2 = dot3(n, a2);
  const float t1_d0 = dot3(n, b0), t1_d1 = dot3(n, b1), t1_d2 = dot3(n, b2);
  const float t0_min = smin(smin(t0_d0, t0_d1), t0_d2);
  const float t1_min = smin(smin(t1_d0, t1_d1), t1_d2);
  const float t0_max = smax(smax(t0_d0, t0_d1), t0_d2);
  const float t1_max = smax(smax(t1_d0, t1_d1), t1_d2);
  return (t0_max < t1_min) || (t0_min > t1_max);
}

// Candidate provenance, one byte: kind (bits 0-1), vertex or edge (bits 2-4),
// plane (bits 5-6).  kind 0: vertex of B inside A; 1: vertex of A inside B;
// 2: edge of B crossing a face plane of A; 3: edge of A crossing a face plane of B.
__device__ inline uint8_t cand_code(int kind, int ve, int f) { return (uint8_t)(kind | (ve << 2) | (f << 5)); }

constexpr int kMaxCand = 56;  // 8 vertices + 6 edges x 4 planes x 2 directions

// Per-wave LDS of the narrow phase: the two tets (A: v[0..3], B: v[4..7]) and
// their face planes (A faces u[0..3], B faces u[4..7]).
struct NarrowLds {
  f3 v[8];
  UPlane u[8];
};

// SAT axis `ax` (0..43) in the reference's order (overlap ...tetrahedron.h:
// 50-290): 0-3 faces of A against B's vertices, 4-7 faces of B against A's,
// 8-43 edge pairs (ea, eb) = ((ax-8)/6, (ax-8)%6).  V: A's vertices, then
// B's.  True when the axis separates.
__device__ inline bool sat_axis(const f3* V, uint32_t ax) {
  const f3* A = V;
  const f3* B = V + 4;
  if (ax < 8u) {
    const f3* P = ax < 4u ? A : B;
    const f3* Q = ax < 4u ? B : A;
    const uint32_t f = ax & 3u;
    f3 n; float w;
    make_plane(P[face_v(f, 0)], P[face_v(f, 1)], P[face_v(f, 2)], n, w);
    const float d0 = dot3(n, Q[0]) - w, d1 = dot3(n, Q[1]) - w, d2 = dot3(n, Q[2]) - w, d3 = dot3(n, Q[3]) - w;
    return d0 > 0.0f && d1 > 0.0f && d2 > 0.0f && d3 > 0.0f;
  }
  const uint32_t k = ax - 8u, ea = k / 6u, eb = k - 6u * ea;
  const f3 eA = sub3(A[edge_v(ea, 1)], A[edge_v(ea, 0)]);
  const f3 eB = sub3(B[edge_v(eb, 1)], B[edge_v(eb, 0)]);
  const f3 n = cross3(eA, eB);
  const uint32_t ta = sat_a(ea), tb = sat_b(ea, eb);
  return sat_line(A[triple_v(ta, 0)], A[triple_v(ta, 1)], A[triple_v(ta, 2)], B[triple_v(tb, 0)], B[triple_v(tb, 1)],
                  B[triple_v(tb, 2)], n);
}

// Contact candidate `c` (0..55) in the reference's generation order
// (contacts ...tetrahedron.h:543-640): c < 8: vertex c/2, B's inside A (even c)
// or A's inside B (odd c); c >= 8: k = c - 8, edge k/8 against face plane
// (k/2)%4, B's edge on A's plane (even k) or A's edge on B's plane (odd k).
// Returns whether it exists; *q its point, *code its provenance.
__device__ inline bool candidate(const NarrowLds& L, uint32_t c, f3* q, uint8_t* code) {
  const f3* A = L.v;
  const f3* B = L.v + 4;
  const UPlane* PA = L.u;
  const UPlane* PB = L.u + 4;
  if (c < 8u) {
    const uint32_t v = c >> 1;
    if (c & 1u) {
      *q = A[v];
      *code = cand_code(1, (int)v, 0);
      return inside_planes(*q, PB);
    }
    *q = B[v];
    *code = cand_code(0, (int)v, 0);
    return inside_planes(*q, PA);
  }
  const uint32_t k = c - 8u, e = k >> 3, p = (k >> 1) & 3u;
  const uint32_t i = edge_v(e, 0), j = edge_v(e, 1);
  const bool a_edge = k & 1u;
  const f3* E = a_edge ? A : B;
  const UPlane* P = a_edge ? PB : PA;
  *code = cand_code(a_edge ? 3 : 2, (int)e, (int)p);
  if (!is_crossing(E[i], E[j], P[p])) return false;
  *q = make_intersection(E[i], E[j], P[p]);
  return inside_planes(*q, P);
}

}  // namespace kdop_amd
