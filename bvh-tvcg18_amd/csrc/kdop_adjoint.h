// kdop_adjoint.h -- adjoint of the narrow-phase penetration measure.
//
// Phi(pair) = sum over the tet pairs of that body pair that reported contacts
// of nuniq * depth, i.e. the sum of the depths of all its reported contacts.
// The discrete choices of the forward narrow phase are frozen (the A face that
// gives the normal, and the provenance of the candidates attaining the minimum
// and maximum of p.n, recorded per tet pair in NarrowRec), so Phi is a smooth
// function of the two bodies' poses:
//   depth = (p_min - p_max) . n,  n = unit normal of the frozen face of A,
//   p = a vertex, or the crossing of an edge with a face plane at the signed
//       parameter t = sA / (sA - sB) (equal to the forward's |sA| / (|sA|+|sB|)
//       wherever the edge crosses, and smooth where that form has a kink),
//   world vertex r = Im(q r0 q*) + p (the literal transform, q not normalised,
//   tiny_quaternion_functions.h:137-152 as narrow_update_kdop_bvh.h:154-165
//   applies it).
// Everything here runs in FP64 from the material coordinates and the poses of
// the last refit.  The reference has no differentiable narrow phase.
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

struct d3 {
  double x, y, z;
};
__device__ inline d3 dmk(double a, double b, double c) { return d3{a, b, c}; }
__device__ inline double ddot(const d3& a, const d3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline d3 dcross(const d3& a, const d3& b) { return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline d3 dsub(const d3& a, const d3& b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline d3 dadd(const d3& a, const d3& b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline d3 dmul(const d3& a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }

// r = Im(q r0 q*) = (s^2 - u.u) r0 + 2 (u.r0) u + 2 s (u x r0)
__device__ inline d3 drotate(const double* q, const d3& r0) {
  const d3 u = dmk(q[0], q[1], q[2]);
  const double s = q[3];
  return dadd(dadd(dmul(r0, s * s - ddot(u, u)), dmul(u, 2.0 * ddot(u, r0))), dmul(dcross(u, r0), 2.0 * s));
}

// Plane of face f of tet P: N = (p1 - p0) x (p2 - p0) through p0, the vertices
// of the face opposite vertex f in the reference's order.
struct DPlane {
  d3 N, p0;
  uint32_t f0, f1, f2;
};
__device__ inline DPlane dface(const d3* P, uint32_t f) {
  DPlane pl;
  pl.f0 = face_v(f, 0); pl.f1 = face_v(f, 1); pl.f2 = face_v(f, 2);
  pl.N = dcross(dsub(P[pl.f1], P[pl.f0]), dsub(P[pl.f2], P[pl.f0]));
  pl.p0 = P[pl.f0];
  return pl;
}

// Candidate point of provenance `code` (kind | vertex-or-edge << 2 | plane << 5).
__device__ inline d3 dcand(const d3* A, const d3* B, uint32_t code) {
  const uint32_t kind = code & 3u, ve = (code >> 2) & 7u, f = (code >> 5) & 3u;
  if (kind == 0u) return B[ve];
  if (kind == 1u) return A[ve];
  const d3* E = kind == 2u ? B : A;  // tet owning the edge
  const d3* F = kind == 2u ? A : B;  // tet owning the plane
  const DPlane pl = dface(F, f);
  const d3 a = E[edge_v(ve, 0)], b = E[edge_v(ve, 1)];
  const double sA = ddot(pl.N, dsub(a, pl.p0)), sB = ddot(pl.N, dsub(b, pl.p0));
  const double t = sA / (sA - sB);
  return dadd(a, dmul(dsub(b, a), t));
}

// d/d(p0,p1,p2) of N = (p1 - p0) x (p2 - p0) contracted with gN
__device__ inline void dgrad_plane(const d3* P, const DPlane& pl, const d3& gN, d3* g) {
  const d3 e1 = dsub(P[pl.f1], P[pl.f0]), e2 = dsub(P[pl.f2], P[pl.f0]);
  const d3 ge1 = dcross(e2, gN), ge2 = dcross(gN, e1);
  g[pl.f1] = dadd(g[pl.f1], ge1);
  g[pl.f2] = dadd(g[pl.f2], ge2);
  g[pl.f0] = dsub(g[pl.f0], dadd(ge1, ge2));
}

// Reverse mode of one candidate point, adjoint gq.
__device__ inline void dgrad_cand(const d3* A, const d3* B, uint32_t code, const d3& gq, d3* gA, d3* gB) {
  const uint32_t kind = code & 3u, ve = (code >> 2) & 7u, f = (code >> 5) & 3u;
  if (kind == 0u) { gB[ve] = dadd(gB[ve], gq); return; }
  if (kind == 1u) { gA[ve] = dadd(gA[ve], gq); return; }
  const d3* E = kind == 2u ? B : A;
  const d3* F = kind == 2u ? A : B;
  d3* gE = kind == 2u ? gB : gA;
  d3* gF = kind == 2u ? gA : gB;
  const uint32_t ia = edge_v(ve, 0), ib = edge_v(ve, 1);
  const d3 a = E[ia], b = E[ib];
  const DPlane pl = dface(F, f);
  const double sA = ddot(pl.N, dsub(a, pl.p0)), sB = ddot(pl.N, dsub(b, pl.p0));
  const double den = sA - sB, t = sA / den;
  gE[ia] = dadd(gE[ia], dmul(gq, 1.0 - t));
  gE[ib] = dadd(gE[ib], dmul(gq, t));
  const double gt = ddot(dsub(b, a), gq);
  const double gsA = gt * (-sB / (den * den)), gsB = gt * (sA / (den * den));
  gE[ia] = dadd(gE[ia], dmul(pl.N, gsA));
  gE[ib] = dadd(gE[ib], dmul(pl.N, gsB));
  const d3 gN = dadd(dmul(dsub(a, pl.p0), gsA), dmul(dsub(b, pl.p0), gsB));
  gF[pl.f0] = dsub(gF[pl.f0], dmul(pl.N, gsA + gsB));
  dgrad_plane(F, pl, gN, gF);
}

// Frozen depth and its gradient (weight w) with respect to the 8 vertices.
__device__ inline double dfrozen_depth_grad(const d3* A, const d3* B, uint32_t face, uint32_t cmin, uint32_t cmax, double w, d3* gA,
                                            d3* gB) {
  const DPlane pl = dface(A, face);
  const double len = sqrt(ddot(pl.N, pl.N));
  const d3 n = dmul(pl.N, 1.0 / len);
  const d3 pmin = dcand(A, B, cmin), pmax = dcand(A, B, cmax);
  const double depth = ddot(pmin, n) - ddot(pmax, n);
  dgrad_cand(A, B, cmin, dmul(n, w), gA, gB);
  dgrad_cand(A, B, cmax, dmul(n, -w), gA, gB);
  const d3 gn = dmul(dsub(pmin, pmax), w);
  const d3 gN = dmul(dsub(gn, dmul(n, ddot(n, gn))), 1.0 / len);
  dgrad_plane(A, pl, gN, gA);
  return depth;
}

// Pose gradient of r = (s^2 - u.u) r0 + 2 (u.r0) u + 2 s (u x r0) + p: adds to
// g[0..3) (position) and g[3..7) (quaternion x y z w).
__device__ inline void dpose_grad(const double* q, const d3& r0, const d3& gr, double* g) {
  const d3 u = dmk(q[0], q[1], q[2]);
  const double s = q[3];
  g[0] += gr.x; g[1] += gr.y; g[2] += gr.z;
  const d3 gu = dadd(dadd(dmul(u, -2.0 * ddot(gr, r0)), dmul(r0, 2.0 * ddot(gr, u))),
                     dadd(dmul(gr, 2.0 * ddot(u, r0)), dmul(dcross(r0, gr), 2.0 * s)));
  g[3] += gu.x; g[4] += gu.y; g[5] += gu.z;
  g[6] += ddot(gr, dadd(dmul(r0, 2.0 * s), dmul(dcross(u, r0), 2.0)));
}

constexpr int kAdjStripes = 16;  // accumulator stripes per body pair
constexpr int kAdjVals = 16;     // per stripe: phi, 7 pose grads of A, 7 of B, pad

struct AdjArgs {
  const BodyDev* bodies;
  const uint2* pairs;
  const NarrowRec* recs;
  uint32_t nrec;
  const float* poses;  // 8 floats per body
  const float* dphi;   // per body pair, or null (all ones)
  double* acc;         // [npairs][kAdjStripes][kAdjVals]
};

// One thread per narrow record: frozen depth and its pose gradients in FP64,
// accumulated per body pair (one atomic per wave and value when the wave's
// records belong to one pair, the usual case since a pair's records are
// written together).
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs A) {
  const uint32_t n = A.nrec;
  const uint32_t lane = lane_id();
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    double v[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    uint32_t pair = 0xFFFFFFFFu;
    if (i < n) {
      const NarrowRec r = A.recs[i];
      pair = r.pair;
      const uint2 bp = A.pairs[pair];
      const BodyDev& BA = A.bodies[bp.x];
      const BodyDev& BB = A.bodies[bp.y];
      double qa[4], qb[4];
      d3 pa, pb;
      {
        const float* a = A.poses + 8 * (size_t)bp.x;
        const float* b = A.poses + 8 * (size_t)bp.y;
        pa = dmk(a[0], a[1], a[2]);
        pb = dmk(b[0], b[1], b[2]);
        for (int k = 0; k < 4; ++k) { qa[k] = a[3 + k]; qb[k] = b[3 + k]; }
      }
      d3 A0[4], B0[4], Aw[4], Bw[4], gA[4], gB[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t ia = BA.tets[4 * (size_t)r.ta + k], ib = BB.tets[4 * (size_t)r.tb + k];
        A0[k] = dmk(BA.X0[ia], BA.Y0[ia], BA.Z0[ia]);
        B0[k] = dmk(BB.X0[ib], BB.Y0[ib], BB.Z0[ib]);
        Aw[k] = dadd(drotate(qa, A0[k]), pa);
        Bw[k] = dadd(drotate(qb, B0[k]), pb);
        gA[k] = dmk(0.0, 0.0, 0.0);
        gB[k] = dmk(0.0, 0.0, 0.0);
      }
      const double w = (double)r.nuniq * (A.dphi ? (double)A.dphi[pair] : 1.0);
      const double depth = dfrozen_depth_grad(Aw, Bw, r.face, r.cmin, r.cmax, w, gA, gB);
      v[0] = (double)r.nuniq * depth;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dpose_grad(qa, A0[k], gA[k], v + 1);
        dpose_grad(qb, B0[k], gB[k], v + 8);
      }
    }
    // the wave's records of one pair: reduce first, one atomic per value
    // lanes past the end have no record; lane 0 has one unless the whole wave is past the end
    const uint32_t p0 = __builtin_amdgcn_readfirstlane(pair);
    const bool one = __all(pair == p0 || pair == 0xFFFFFFFFu) && p0 != 0xFFFFFFFFu;
    const uint32_t stripe = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (kAdjStripes - 1);
    if (one) {
#pragma unroll
      for (int k = 0; k < 15; ++k) {
        double x = v[k];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) atomicAdd(A.acc + ((size_t)p0 * kAdjStripes + stripe) * kAdjVals + k, x);
      }
    } else if (pair != 0xFFFFFFFFu) {
      for (int k = 0; k < 15; ++k) atomicAdd(A.acc + ((size_t)pair * kAdjStripes + stripe) * kAdjVals + k, v[k]);
    }
  }
}

// Per requested body: the sum over the call's body pairs it belongs to (as A
// or B) of the accumulated pose gradient; per pair: Phi.  One thread per
// (body, component) and per pair.
__global__ __launch_bounds__(256) void k_adjoint_finish(const uint2* __restrict__ pairs, uint32_t npairs, const double* __restrict__ acc,
                                                        const uint32_t* __restrict__ bodies, uint32_t nbodies, float* __restrict__ grad_pos,
                                                        float* __restrict__ grad_quat, float* __restrict__ phi) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < 7u * nbodies) {
    const uint32_t b = g / 7u, c = g - 7u * b, body = bodies[b];
    double s = 0.0;
    for (uint32_t p = 0; p < npairs; ++p) {
      const uint2 bp = pairs[p];
      for (int half = 0; half < 2; ++half) {
        if ((half == 0 ? bp.x : bp.y) != body) continue;
        for (int st = 0; st < kAdjStripes; ++st) s += acc[((size_t)p * kAdjStripes + st) * kAdjVals + 1 + 7 * half + c];
      }
    }
    if (c < 3u) grad_pos[3 * (size_t)b + c] = (float)s;
    else grad_quat[4 * (size_t)b + (c - 3u)] = (float)s;
  }
  if (phi && g < npairs) {
    double s = 0.0;
    for (int st = 0; st < kAdjStripes; ++st) s += acc[((size_t)g * kAdjStripes + st) * kAdjVals];
    phi[g] = (float)s;
  }
}

}  // namespace kdop_amd
