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

// faces opposite vertex 0..3 (types/geometry_tetrahedron.h:231-257)
__constant__ static const int8_t kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
// edges 01 02 03 12 13 23 (overlap ...tetrahedron.h:102-107)
__constant__ static const int8_t kEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
// Vertex triples SAT_line projects for each edge-edge axis, as the reference
// hard-codes them (overlap ...tetrahedron.h:109-287): A side by A edge, B side
// by (A edge, B edge); triple ids 0={1,2,3} 1={0,1,3} 2={0,1,2}.
__constant__ static const int8_t kTriple[3][3] = {{1, 2, 3}, {0, 1, 3}, {0, 1, 2}};
__constant__ static const int8_t kSatA[6] = {0, 0, 0, 1, 2, 2};
__constant__ static const int8_t kSatB[6][6] = {{0, 1, 2, 1, 2, 2}, {0, 0, 2, 1, 2, 2}, {0, 0, 0, 1, 2, 2},
                                                 {0, 0, 0, 1, 2, 2}, {0, 0, 0, 1, 2, 2}, {0, 0, 0, 1, 2, 2}};

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
__device__ inline UPlane face_uplane(const f3* t, int f) { return make_uplane(t[kFace[f][0]], t[kFace[f][1]], t[kFace[f][2]]); }
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

// overlap/geometry_overlap_tetrahedron_tetrahedron.h:50-290
__device__ inline bool overlap_tet_tet(const f3* A, const f3* B) {
  for (int i = 0; i < 4; ++i) {
    f3 n; float w;
    make_plane(A[kFace[i][0]], A[kFace[i][1]], A[kFace[i][2]], n, w);
    const float b0 = dot3(n, B[0]) - w, b1 = dot3(n, B[1]) - w, b2 = dot3(n, B[2]) - w, b3 = dot3(n, B[3]) - w;
    if (b0 > 0.0f && b1 > 0.0f && b2 > 0.0f && b3 > 0.0f) return false;
  }
  for (int i = 0; i < 4; ++i) {
    f3 n; float w;
    make_plane(B[kFace[i][0]], B[kFace[i][1]], B[kFace[i][2]], n, w);
    const float a0 = dot3(n, A[0]) - w, a1 = dot3(n, A[1]) - w, a2 = dot3(n, A[2]) - w, a3 = dot3(n, A[3]) - w;
    if (a0 > 0.0f && a1 > 0.0f && a2 > 0.0f && a3 > 0.0f) return false;
  }
  for (int ea = 0; ea < 6; ++ea) {
    const f3 eA = sub3(A[kEdge[ea][1]], A[kEdge[ea][0]]);
    const int8_t* ta = kTriple[kSatA[ea]];
    for (int eb = 0; eb < 6; ++eb) {
      const f3 eB = sub3(B[kEdge[eb][1]], B[kEdge[eb][0]]);
      const f3 n = cross3(eA, eB);
      const int8_t* tb = kTriple[kSatB[ea][eb]];
      if (sat_line(A[ta[0]], A[ta[1]], A[ta[2]], B[tb[0]], B[tb[1]], B[tb[2]], n)) return false;
    }
  }
  return true;
}

// contacts ...tetrahedron.h:464-533; returns the chosen face of A or -1
__device__ inline int pick_most_opposing(const f3* A, const f3* B, uint32_t sa, uint32_t sb, f3& n) {
  int found = -1;
  float best = 0.0f;
  for (int a = 0; a < 4; ++a) {
    if (!((sa >> a) & 1u)) continue;
    const UPlane plA = face_uplane(A, a);
    const f3 nA = unit3(plA.n);
    for (int b = 0; b < 4; ++b) {
      if (!((sb >> b) & 1u)) continue;
      const UPlane plB = face_uplane(B, b);
      const f3 nB = unit3(plB.n);
      const float test = dot3(nA, nB);
      if (test >= 0.0f) continue;
      const float testA = dot3(plA.n, sub3(B[b], plA.p));
      const float testB = dot3(plB.n, sub3(A[a], plB.p));
      if (testA < 0.0f) continue;
      if (testB < 0.0f) continue;
      if (test < best) {
        n = nA;
        best = test;
        found = a;
      }
    }
  }
  return found;
}

// Candidate provenance, one byte: kind (bits 0-1), vertex or edge (bits 2-4),
// plane (bits 5-6).  kind 0: vertex of B inside A; 1: vertex of A inside B;
// 2: edge of B crossing a face plane of A; 3: edge of A crossing a face plane of B.
__device__ inline uint8_t cand_code(int kind, int ve, int f) { return (uint8_t)(kind | (ve << 2) | (f << 5)); }

constexpr int kMaxCand = 56;  // 8 vertices + 6 edges x 4 planes x 2 directions

struct NarrowResult {
  int face;        // A face giving the normal, -1 when no contact
  int ncand;       // candidates generated
  int nuniq;       // reported contacts
  f3 n;
  float depth;
  uint8_t cmin, cmax;  // provenance of the candidates attaining min / max of p.n
};

// contacts ...tetrahedron.h:543-681 + :1312-1345.  On return raw[0..ncand)
// holds the projected candidates and bit c of *mask is set when candidate c is
// reported (not within working precision of an earlier one).
__device__ inline NarrowResult narrow_pair(const f3* A, const f3* B, uint32_t sa, uint32_t sb, f3* raw, uint64_t* mask) {
  NarrowResult r;
  r.face = -1; r.ncand = 0; r.nuniq = 0; r.depth = 0.0f; r.n = f3{0.f, 0.f, 0.f}; r.cmin = r.cmax = 0;
  *mask = 0;
  if (!overlap_tet_tet(A, B)) return r;
  f3 n{0.f, 0.f, 0.f};
  const int fa = pick_most_opposing(A, B, sa, sb, n);
  if (fa < 0) return r;
  UPlane PA[4], PB[4];
  for (int v = 0; v < 4; ++v) { PA[v] = face_uplane(A, v); PB[v] = face_uplane(B, v); }
  uint8_t code[kMaxCand];
  int nc = 0;
  for (int v = 0; v < 4; ++v) {
    if (inside_planes(B[v], PA)) { raw[nc] = B[v]; code[nc++] = cand_code(0, v, 0); }
    if (inside_planes(A[v], PB)) { raw[nc] = A[v]; code[nc++] = cand_code(1, v, 0); }
  }
  for (int e = 0; e < 6; ++e) {
    const int i = kEdge[e][0], j = kEdge[e][1];
    for (int p = 0; p < 4; ++p) {
      if (is_crossing(B[i], B[j], PA[p])) {
        const f3 q = make_intersection(B[i], B[j], PA[p]);
        if (inside_planes(q, PA)) { raw[nc] = q; code[nc++] = cand_code(2, e, p); }
      }
      if (is_crossing(A[i], A[j], PB[p])) {
        const f3 q = make_intersection(A[i], A[j], PB[p]);
        if (inside_planes(q, PB)) { raw[nc] = q; code[nc++] = cand_code(3, e, p); }
      }
    }
  }
  r.face = fa;
  r.n = n;
  r.ncand = nc;
  if (nc == 0) { r.face = -1; return r; }
  float mn = FLT_MAX, mx = -FLT_MAX;
  int imin = 0, imax = 0;
  for (int c = 0; c < nc; ++c) {
    const float d = dot3(raw[c], n);
    if (d < mn) { mn = d; imin = c; }
    if (mx < d) { mx = d; imax = c; }
  }
  r.depth = mn - mx;
  r.cmin = code[imin];
  r.cmax = code[imax];
  const f3 mid = mul3(mul3(n, mx + mn), 0.5f);
  const float wp = FLT_EPSILON * 10.0f;
  int nu = 0;
  uint64_t m = 0;
  for (int c = 0; c < nc; ++c) {
    const float s = dot3(n, sub3(raw[c], mid));
    raw[c] = sub3(raw[c], f3{s * n.x, s * n.y, s * n.z});
    bool uniq = true;
    for (int q = 0; q < c; ++q) {
      const f3 d = sub3(raw[q], raw[c]);
      const float n1 = smax(smax(fabsf(d.x), fabsf(d.y)), fabsf(d.z));
      if (n1 < wp) { uniq = false; break; }
    }
    if (uniq) { m |= 1ull << c; ++nu; }
  }
  *mask = m;
  r.nuniq = nu;
  return r;
}

}  // namespace kdop_amd
