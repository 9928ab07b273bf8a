// kdop_kernels.h -- HIP kernels of the kDOP pipeline (build, refit, traversal,
// narrow phase).  Included once, by kdop_hip.hip.
#pragma once

#include "kdop_device.h"
#include "kdop_math.h"
#include "kdop_narrow.h"
#include "kdop_amd.h"

namespace kdop_amd {

// ------------------------------------------------------------- wave helpers
__device__ inline int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, (unsigned)d, 64);
    if (lane >= d) v += t;
  }
  return v;
}
__device__ inline unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
// Reserve `cnt` slots per lane in a buffer filled through *ctr (one atomic per
// wave).  Every lane of the wave must call it.  Returns this lane's first slot.
__device__ inline uint32_t wave_reserve(uint32_t* ctr, uint32_t cnt) {
  const uint32_t incl = wave_incl_scan(cnt);
  const uint32_t total = __shfl(incl, 63, 64);
  uint32_t base = 0;
  if (lane_id() == 0 && total) base = atomicAdd(ctr, total);
  base = __shfl(base, 0, 64);
  return base + incl - cnt;
}
__device__ inline void wave_add64(unsigned long long* dst, unsigned long long v) {
  const unsigned long long s = wave_sum64(v);
  if (lane_id() == 0 && s) atomicAdd(dst, s);
}
// Exclusive prefix sum across the wave of a per-lane count below 8, from
// three ballots (no LDS round trips); *total receives the wave sum.
__device__ inline uint32_t wave_prefix_small(uint32_t v, uint32_t* total) {
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const unsigned long long m = __ballot((v >> b) & 1u);
    pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
    tot += (uint32_t)__popcll(m) << b;
  }
  *total = tot;
  return pre;
}
// Value of lane 0 (a scalar register read, not an LDS permute).
__device__ inline uint32_t lane0(uint32_t v) { return __builtin_amdgcn_readlane(v, 0); }
// A value the caller knows to be equal on every lane, as a scalar.
__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Orders one wave's LDS accesses: the LDS executes a wave's instructions in
// issue order, so only the compiler must be kept from reordering them (no
// barrier: other waves of the workgroup are not involved).
__device__ inline void wave_lds_sync() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// Global index of the calling wave.
__device__ inline uint32_t wave_global() { return (blockIdx.x * blockDim.x + threadIdx.x) >> 6; }

// Reserve `cnt` (< 8) slots per thread for the whole block with one atomic
// (every thread of the block must call it; lds: nwaves + 1 words).  Returns
// this thread's first slot.
__device__ inline uint32_t block_reserve(uint32_t* ctr, uint32_t cnt, uint32_t* lds) {
  uint32_t wtot;
  const uint32_t incl = wave_prefix_small(cnt, &wtot) + cnt;
  const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane_id() == 0) lds[w] = wtot;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t i = 0; i < nw; ++i) {
      const uint32_t t = lds[i];
      lds[i] = acc;
      acc += t;
    }
    lds[nw] = acc ? atomicAdd(ctr, acc) : 0u;
  }
  __syncthreads();
  const uint32_t r = lds[nw] + lds[w] + incl - cnt;
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ node records
template <int K>
struct NodeRec {
  float v[K];
  uint32_t m0, m1;
};

// Load the used part of a record (K + 2 floats) with 16-byte loads; the
// record stride is a multiple of 4 floats, so the last load stays inside it.
template <int K>
__device__ inline void load_rec(const float* __restrict__ nodes, uint32_t idx, NodeRec<K>& r) {
  constexpr int S = rec_stride(K);
  constexpr int NV = (K + 2 + 3) / 4;
  const float4* q = reinterpret_cast<const float4*>(nodes + (size_t)idx * S);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const float4 x = q[i];
    const float t[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int f = 4 * i + j;
      if (f < K) r.v[f] = t[j];
      else if (f == K) r.m0 = __float_as_uint(t[j]);
      else if (f == K + 1) r.m1 = __float_as_uint(t[j]);
    }
  }
}
template <int K>
__device__ inline uint2 load_meta(const float* __restrict__ nodes, uint32_t idx) {
  return *reinterpret_cast<const uint2*>(nodes + (size_t)idx * rec_stride(K) + K);
}

// geometry_overlap_dop_dop.h:12-24 with the strict interval test of
// geometry_overlap_interval_interval.h:11-21 (no early exit; same result).
template <int K>
__device__ inline bool dop_overlap(const float* A, const float* B) {
  bool sep = false;
#pragma unroll
  for (int k = 0; k < K / 2; ++k) sep |= (A[2 * k + 1] < B[2 * k]) | (B[2 * k + 1] < A[2 * k]);
  return !sep;
}

// ============================================================== build
// One thread per chunk (all levels of one body): writes the meta words of the
// chunk's nodes in BFS order of the reference's median split
// (kdop_make_tree.h:57-121 branch chunks, :123-186 super chunks).
__device__ inline uint32_t chunk_elems(const TopoParams& P, int h, uint32_t c) {
  return (c + 1 < P.C[h]) ? P.L : P.E[h] - c * P.L;
}
__device__ inline uint32_t chunk_stride(const TopoParams& P, int h) { return (h == (int)P.H - 1) ? 2 * P.L : P.L; }
// Root node of chunk (h, c) after skipping one-element super chunks.
__device__ inline uint32_t resolved_root(const TopoParams& P, int h, uint32_t c) {
  for (;;) {
    const uint32_t n = chunk_elems(P, h, c);
    if (h == (int)P.H - 1 || n >= 2) return P.node_base + P.level_base[h] + c * chunk_stride(P, h) + 1;
    c = c * P.L;
    ++h;
  }
}

__global__ void k_build_topology(TopoParams P, float* __restrict__ nodes, int K, const uint8_t* __restrict__ surf,
                                 uint32_t total_chunks) {
  const int S = rec_stride(K);
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_chunks) return;
  int h = 0;
  uint32_t c = g;
  while (h < (int)P.H && c >= P.C[h]) { c -= P.C[h]; ++h; }
  const bool branch = (h == (int)P.H - 1);
  const uint32_t n = chunk_elems(P, h, c);
  const uint32_t base = P.node_base + P.level_base[h] + c * chunk_stride(P, h) + 1;  // slot of BFS node 0
  const uint32_t first = c * P.L;  // first element (tet or child chunk) of the chunk
  if (!branch && n < 2) return;    // no nodes: parents point at the resolved child
  // BFS over element ranges; branch chunks store every range, super chunks
  // only ranges of two or more elements.
  uint16_t qlo[64], qhi[64];
  int head = 0, tail = 0;
  qlo[tail] = 0; qhi[tail] = (uint16_t)(n - 1); ++tail;
  while (head < tail) {
    const uint32_t lo = qlo[head], hi = qhi[head];
    const uint32_t node = base + head;
    uint32_t* meta = reinterpret_cast<uint32_t*>(nodes + (size_t)node * S + K);
    const uint32_t rootflag = (branch && head == 0) ? F_ROOT : 0u;
    ++head;
    if (lo == hi) {  // branch leaf (super chunks never enqueue single ranges)
      const uint32_t tet = first + lo;
      meta[0] = F_LEAF | F_BRANCH | tet;
      meta[1] = rootflag | ((surf[tet] & 15u) ? F_SURF : 0u) | P.body;
      continue;
    }
    const uint32_t mid = (lo + hi) >> 1;  // floor((first+last)/2): first is even (c*L), see header
    uint32_t child[2];
    const uint32_t rlo[2] = {lo, mid + 1}, rhi[2] = {mid, hi};
    for (int s = 0; s < 2; ++s) {
      if (branch || rlo[s] < rhi[s]) {
        child[s] = base + tail;
        qlo[tail] = (uint16_t)rlo[s]; qhi[tail] = (uint16_t)rhi[s]; ++tail;
      } else {
        child[s] = resolved_root(P, h + 1, first + rlo[s]);
      }
    }
    meta[0] = (branch ? F_BRANCH : 0u) | child[0];
    meta[1] = rootflag | child[1];
  }
}

// ============================================================== refit
// World position of vertex v of B under the pose (q, p): rotate(q, r0) + p
// (narrow_update_kdop_bvh.h:154-165), or r0 itself for the identity, which
// make_tree refits to.  Every kernel that needs a vertex of the pose being
// refitted computes it through this function, so all of them get the bits
// that the refit stores.
__device__ inline f3 posed_vertex(const BodyDev& B, uint32_t v, const float* q, const float* p, bool identity) {
  const float4 w0 = B.W0[v];
  const f3 r0 = f3{w0.x, w0.y, w0.z};
  if (identity) return r0;
  const f3 r = rotate3(q[0], q[1], q[2], q[3], r0);
  return f3{r.x + p[0], r.y + p[1], r.z + p[2]};
}

// (x[i], y[i], z[i], 0) for i < n: the rest positions as one 16-byte word per vertex.
__global__ void k_interleave3(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z, uint32_t n,
                              float4* __restrict__ w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = make_float4(x[i], y[i], z[i], 0.0f);
}

// Store the K volume floats of one record (16-byte stores; K is even, so the
// tail is one 8-byte store and the meta words after it stay untouched).
template <int K>
__device__ inline void store_vol(float* __restrict__ nodes, uint32_t slot, const float* v) {
  float* r = nodes + (size_t)slot * rec_stride(K);
#pragma unroll
  for (int f = 0; f + 4 <= K; f += 4) *reinterpret_cast<float4*>(r + f) = make_float4(v[f], v[f + 1], v[f + 2], v[f + 3]);
  if (K % 4) *reinterpret_cast<float2*>(r + (K & ~3)) = make_float2(v[K - 2], v[K - 1]);
}
// Minimum / maximum of slab bounds built from leaf projections.  A
// projection dot3(d, P) starts from +0 (0 + d.x P.x), so in round-to-nearest
// it is never -0; it is NaN only for a NaN or infinite vertex.  Over such
// values the reference's std::min / std::max ((b < a) ? b : a) and IEEE
// minNum / maxNum (v_min_f32 / v_max_f32, which skip a NaN operand) give the
// same bits: equal values have equal bits, and a NaN projection is skipped by
// both when the running bound (never NaN) is the other operand.  One VALU
// instruction (three-operand forms fold chains) instead of a compare and a
// select.
__device__ inline float bmin(float a, float b) { return fminf(a, b); }
__device__ inline float bmax(float a, float b) { return fmaxf(a, b); }
// Union of two slab bounds in the refit butterflies: max(a, b) for an upper
// bound, min(a, b) for a lower one.  A bound is never NaN (a leaf's starts
// from +-FLT_MAX and bmin / bmax skip NaN projections) and never -0, so the
// median of (a, b, +inf) is max(a, b) and of (a, b, -inf) is min(a, b), the
// reference's std::max / std::min bits: one v_med3_f32.  fminf / fmaxf cost
// a canonicalisation of every operand that comes from a lane move or a load,
// and a lane whose kind is known only at run time would compute both and
// select; here the third operand picks the kind.  When the kind is known at
// compile time the infinity is made opaque (an empty asm statement that
// holds it in an SGPR: no instruction), or the compiler folds the median
// back into a min / max with canonicalised operands.
__device__ inline float ubound(float a, float b, bool upper) { return __builtin_amdgcn_fmed3f(a, b, upper ? INFINITY : -INFINITY); }
template <bool UPPER>
__device__ inline float ubound(float a, float b) {
  float lim = UPPER ? INFINITY : -INFINITY;
  asm("" : "+s"(lim));
  return __builtin_amdgcn_fmed3f(a, b, lim);
}

// Leaf DOP of one tetrahedron (geometry_dop.h:48-74): per slab direction the
// min / max projection of its four world vertices (bmin / bmax, above).
// pos(i): world position of vertex i.
template <int K, class Pos>
__device__ inline void tet_dop_at(const BodyDev& B, uint32_t tet, const DirTable& dt, Pos&& pos, float* v) {
  const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)tet);
  const uint32_t vi[4] = {tv.x, tv.y, tv.z, tv.w};
  f3 P[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) P[p] = pos(vi[p]);
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const f3 d = f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]};
    float lo = FLT_MAX, hi = -FLT_MAX;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float proj = dot3(d, P[p]);
      lo = bmin(lo, proj);
      hi = bmax(hi, proj);
    }
    v[2 * k] = lo;
    v[2 * k + 1] = hi;
  }
}
// The same from the body's stored world vertices.
template <int K>
__device__ inline void tet_dop(const BodyDev& B, uint32_t tet, const DirTable& dt, float* v) {
  tet_dop_at<K>(B, tet, dt, [&](uint32_t i) { return f3{B.X[i], B.Y[i], B.Z[i]}; }, v);
}

// Value of lane (lane ^ S) for S in {1, 2, 4, 8, 16, 32}: DPP row moves within
// 16-lane rows (full-rate VALU, no LDS), a swizzle for 16, a permute for 32.
template <int S>
__device__ inline float lane_xor(float v) {
  const int x = __float_as_int(v);
  if constexpr (S == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
  if constexpr (S == 2) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
  if constexpr (S == 4 || S == 8) {
    const int up = __builtin_amdgcn_mov_dpp(x, 0x100 + S, 0xF, 0xF, false);  // row_shl:S, lane i <- i + S
    const int dn = __builtin_amdgcn_mov_dpp(x, 0x110 + S, 0xF, 0xF, false);  // row_shr:S, lane i <- i - S
    return __int_as_float((threadIdx.x & S) ? dn : up);
  }
  if constexpr (S == 16) return __int_as_float(__builtin_amdgcn_ds_swizzle(x, 0x401F));  // bitmask mode, xor 16 in 32 lanes
  if constexpr (S == 32) return __int_as_float(__builtin_amdgcn_ds_bpermute((int)((threadIdx.x ^ 32u) << 2), x));
  return v;
}

// One butterfly level of the chunk unions: lanes j and j ^ S hold the two
// children and both end with their union.  The union of bounds of leaf
// projections does not depend on the operand order (bmin / bmax), so the
// left child need not come first as in geometry_make_union.h:9-25.
// A lane of the other half of this lane's aligned group of 2S lanes.  When
// the lanes of each aligned group of S hold the same value (a butterfly after
// its first levels), any such lane is as good a partner as lane ^ S, and for
// S = 4 and 8 one DPP mirror reaches one where lane_xor takes two row shifts
// and a select.
template <int S>
__device__ inline float lane_other(float v) {
  const int x = __float_as_int(v);
  if constexpr (S == 4) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false));  // row_half_mirror: i <- 7 - i in 8 lanes
  if constexpr (S == 8) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false));  // row_mirror: i <- 15 - i in 16 lanes
  return lane_xor<S>(v);
}
// One level of the chunk butterfly: level t (S = 2^(t-1)) leaves every aligned
// group of 2S lanes with the union of its lanes' volumes.
template <int K, int S>
__device__ inline void union_level(float* v, uint32_t j) {
  (void)j;
#pragma unroll
  for (int f = 0; f < K; ++f) {
    const float y = lane_other<S>(v[f]);
    v[f] = (f & 1) ? ubound<true>(v[f], y) : ubound<false>(v[f], y);
  }
}

// Level H-2 of a wave's full super chunks (L = 2^LOGL, L * L <= 64): the
// wave's 64 / L chunk roots are in LDS rows 0.. of `st` (row stride RS); lane
// (element e = lane >> LOGL, component group cg = lane & (L-1)) takes the
// components cg, cg + L, ... of element e, so the butterfly over the
// elements (lane xor L, 2L, ...) holds no copies; even components are
// minima, odd ones maxima, and L is even, so a lane does one kind only.  The
// super chunks' nodes go to rows sc * L + BFS index (sact: this lane's super
// chunk is full), overwriting the roots.  Unions by bmin / bmax (their
// operand order does not matter).
template <int K, int LOGL, uint32_t RS>
__device__ inline void super_butterfly(float* __restrict__ st, uint32_t lane, bool sact) {
  constexpr uint32_t L = 1u << LOGL, NC = (K + L - 1) / L;
  const uint32_t e = lane >> LOGL, cg = lane & (L - 1u), sc = e >> LOGL, ei = e & (L - 1u);
  const bool hi = (cg & 1u) != 0u;
  float w[NC];
#pragma unroll
  for (uint32_t m = 0; m < NC; ++m) w[m] = (cg + m * L < (uint32_t)K) ? st[e * RS + cg + m * L] : 0.0f;
  wave_lds_sync();
#pragma unroll
  for (uint32_t t = 1; t <= LOGL; ++t) {
    const uint32_t S = L << (t - 1u);
#pragma unroll
    for (uint32_t m = 0; m < NC; ++m) {
      float y;
      switch (S) {
        case 2: y = lane_xor<2>(w[m]); break;
        case 4: y = lane_xor<4>(w[m]); break;
        case 8: y = lane_xor<8>(w[m]); break;
        case 16: y = lane_xor<16>(w[m]); break;
        default: y = lane_xor<32>(w[m]); break;
      }
      w[m] = ubound(w[m], y, hi);
    }
    if (sact && (ei & ((2u << (t - 1u)) - 1u)) == 0u) {
      float* row = st + (sc * L + (1u << (LOGL - t)) - 1u + (ei >> t)) * RS;
#pragma unroll
      for (uint32_t m = 0; m < NC; ++m)
        if (cg + m * L < (uint32_t)K) row[cg + m * L] = w[m];
    }
  }
}

// Level H-2 of one full super chunk of L = 16 branch chunks in one wave: the
// 16 chunk roots are rows 0..15 of `roots` (row stride RS); lane (element
// e = lane >> 2, component group cg = lane & 3) takes the components cg,
// cg + 4, ... of element e, and the butterfly over the elements (lane xor 4,
// 8, 16, 32) leaves the 15 nodes in rows 0..14 (BFS order) of `st`.  Even
// components are minima, odd ones maxima (bmin / bmax: the operand order does
// not matter).
template <int K, uint32_t RS>
__device__ inline void super16_butterfly(const float* __restrict__ roots, float* __restrict__ st, uint32_t lane) {
  constexpr uint32_t NC = (K + 3) / 4;
  const uint32_t e = lane >> 2, cg = lane & 3u;
  const bool hi = (cg & 1u) != 0u;
  float w[NC];
#pragma unroll
  for (uint32_t m = 0; m < NC; ++m) w[m] = (cg + 4u * m < (uint32_t)K) ? roots[e * RS + cg + 4u * m] : 0.0f;
#pragma unroll
  for (uint32_t t = 1; t <= 4; ++t) {
#pragma unroll
    for (uint32_t m = 0; m < NC; ++m) {
      float y;
      switch (t) {
        case 1: y = lane_xor<4>(w[m]); break;
        case 2: y = lane_xor<8>(w[m]); break;
        case 3: y = lane_xor<16>(w[m]); break;
        default: y = lane_xor<32>(w[m]); break;
      }
      w[m] = ubound(w[m], y, hi);
    }
    if ((e & ((1u << t) - 1u)) == 0u) {
      float* row = st + ((1u << (4u - t)) - 1u + (e >> t)) * RS;
#pragma unroll
      for (uint32_t m = 0; m < NC; ++m)
        if (cg + 4u * m < (uint32_t)K) row[cg + 4u * m] = w[m];
    }
  }
}

// Full branch chunks, one thread per tetrahedron: the leaf DOP, then the
// chunk's internal nodes by a butterfly over its L lanes.  A full chunk's
// median-split tree over L = 2^logL leaves is the implicit heap (BFS node
// 2^d - 1 + i at depth d, children of BFS node i at 2i+1 and 2i+2, leaf j at
// L - 1 + j), so level t of the butterfly is depth logL - t, and every union
// takes (left, right) in the order the reference's bottom-up pass does
// (geometry_make_union.h:9-25).  tpre: prefix over the batch of
// the tets in full chunks, each body's share rounded up to 64 so that a wave
// never mixes bodies (L, a power of two <= 32, then divides every wave's tet
// range and an aligned L-lane group is one chunk).
//
// When L * L <= 64 a wave also holds whole super chunks of level H-2 (L full
// branch chunks each, whose chunk roots are the super chunk's elements; a
// full super chunk is the same heap), and the butterfly goes on over them.
// When L = 16 a block's 256 tets are one super chunk of level H-2 (a body
// with L = 16 has a tet range of its own that starts on a block), so after
// one barrier wave 0 takes the block's 16 chunk roots from LDS and refits
// that chunk (super16_butterfly).  When every body of the batch has L = 8
// the launch takes 512-thread blocks (NW = 8 waves), a block's 512 tets are
// one super chunk of level H-3, and after one barrier wave 0 refits it from
// its waves' level H-2 roots.  Super chunks that are not full, or that take a
// chunk that is not, are left to k_refit_super.
//
// The volumes are staged in LDS (the leaves, the branch chunks' internal
// nodes, then the super chunks' nodes; each a contiguous run of records) and
// written out with consecutive 16-byte words on consecutive lanes, so that a
// store instruction covers whole lines instead of one word of 64 lines.  The
// branch chunks' records are written whole: the DOP, the two meta words and
// the padding, so that every line is written in full.  HBM3E has no write
// masks, and a line written in part costs a read-modify-write beyond the L2
// (104 of 128 bytes per line: 30.8 us for 1.28M records against 25.2 us for
// the whole lines).  The meta words of a full chunk follow from its heap
// layout (k_build_topology wrote the same words): internal node i of a branch
// chunk holds F_BRANCH | its first child (BFS 2i + 1) and its second child
// (BFS 2i + 2, F_ROOT on node 0); leaf j holds F_LEAF | F_BRANCH | its tet and
// its surface flag and body (F_ROOT when L = 1); node i of a super chunk
// holds its children, BFS 2i + 1 and 2i + 2 when they are below L - 1, else
// the root of branch chunk 2i + 2 - L (or 2i + 3 - L) of the super chunk.  The
// padding keeps the 0xFF bytes the node pool is filled with at kdop_add_body.
//
// The same launch is the vertex transform of the batch: thread g also writes
// world vertex g - vpre[item] of its item (vpre: prefix of the vertex counts,
// each rounded up to 64) to the body's X, Y, Z and W, which later kernels
// read, and the first thread of each body records its pose in body_pose (8
// floats per body: p, q as x y z w, 0, for the adjoint).  A leaf does not
// wait for those stores: it computes its four vertices from the rest
// positions with the same arithmetic (posed_vertex).  Every item's vertex and
// tet ranges are whole waves, so a wave finds its items in the per-wave
// tables vwave and twave with one load at a wave-uniform index, and the
// body's descriptor with scalar loads (a binary search over the prefix sums
// per thread is a chain of log2(batch) dependent loads: the refit of 128
// bodies took 512 us with it and takes 379 us with the tables).
// pos / quat: 3 and 4 floats per batch item (unused with identity).
template <int K, int NW>
__global__ __launch_bounds__(64 * NW) void k_refit_leaves(const BodyDev* __restrict__ bodies, const uint32_t* __restrict__ batch_body,
                                                      const uint32_t* __restrict__ tpre, const uint32_t* __restrict__ vpre,
                                                      const uint32_t* __restrict__ twave, const uint32_t* __restrict__ vwave, uint32_t nbatch,
                                                      const float* __restrict__ pos, const float* __restrict__ quat, int identity,
                                                      float* __restrict__ body_pose, float* __restrict__ nodes, DirTable dt) {
  constexpr uint32_t S = rec_stride(K), RS = (K + 3) & ~3, W = S / 4;  // RS: LDS row, W: 16-byte words per record
  static_assert(K % 4 == 0 || K % 4 == 2, "the meta words start a 16-byte word or end one");
  static_assert(NW == 4 || NW == 8, "256- or 512-thread blocks");
  __shared__ __align__(16) float stage[NW][64 * RS];
  __shared__ __align__(16) float roots16[16 * RS];  // the block's chunk roots (L = 16: 16 of level H-1, L = 8: 8 of level H-2)
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t gw = __builtin_amdgcn_readfirstlane(g >> 6);  // this thread's wave of the batch
  if (gw < (vpre[nbatch] >> 6)) {
    const uint32_t vi = vwave[gw];
    const uint32_t bi = batch_body[vi];
    const BodyDev& B = bodies[bi];
    const uint32_t v = g - vpre[vi];
    const float* q = quat + 4 * (size_t)vi;
    const float* p = pos + 3 * (size_t)vi;
    if (v == 0u) {
      float* bp = body_pose + 8 * (size_t)bi;
      if (identity) {
        bp[0] = bp[1] = bp[2] = 0.f;
        bp[3] = bp[4] = bp[5] = 0.f;
        bp[6] = 1.f;
      } else {
        bp[0] = p[0]; bp[1] = p[1]; bp[2] = p[2];
        bp[3] = q[0]; bp[4] = q[1]; bp[5] = q[2]; bp[6] = q[3];
      }
      bp[7] = 0.f;
    }
    if (v < B.nv) {
      const f3 r = posed_vertex(B, v, q, p, identity != 0);
      B.X[v] = r.x; B.Y[v] = r.y; B.Z[v] = r.z;
      B.W[v] = make_float4(r.x, r.y, r.z, 0.0f);
    }
  }
  if (gw >= (tpre[nbatch] >> 6)) return;  // whole waves (tpre[nbatch] is a multiple of 64)
  float* st = stage[threadIdx.x >> 6];  // this wave's rows (the block's roots for its last flush at L = 8)
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t item = twave[gw];
  const BodyDev& B = bodies[batch_body[item]];
  const float* iq = quat + 4 * (size_t)item;
  const float* ip = pos + 3 * (size_t)item;
  const uint32_t logL = (uint32_t)(__ffs((int)B.L) - 1), L = B.L;
  const uint32_t tet = g - tpre[item];
  const uint32_t tet0 = tet - lane;                   // first tet of the wave
  const uint32_t nfull = (B.nt >> logL) << logL;      // tets in full chunks
  const bool act = tet < nfull;                       // whole L-lane groups are active or not
  const uint32_t j = tet & (L - 1u), q = lane >> logL;  // leaf of the chunk, chunk of the wave
  // level H-2 in this wave: super chunk lane >> 2logL covers tets [tet & ~(L*L-1), +L*L)
  const bool fuse = B.H >= 2u && logL >= 1u && logL <= 3u;
  const uint32_t logLL = 2u * logL;
  const bool sact = fuse && (tet >> logLL << logLL) + (1u << logLL) <= nfull;
  float* __restrict__ chunk_nodes = nodes + ((size_t)B.branch_base + (size_t)(tet0 >> logL) * (2u * L) + 1u) * S;  // BFS node 0 of chunk 0
  float v[K];
#pragma unroll
  for (int f = 0; f < K; ++f) v[f] = 0.0f;
  auto stage_row = [&](uint32_t row) {
    float* r = st + row * RS;
#pragma unroll
    for (int f = 0; f + 4 <= K; f += 4) *reinterpret_cast<float4*>(r + f) = make_float4(v[f], v[f + 1], v[f + 2], v[f + 3]);
    if (K % 4) *reinterpret_cast<float2*>(r + (K & ~3)) = make_float2(v[K - 2], v[K - 1]);
  };
  // staged rows [0, nrows) -> whole records: the DOP from the row, the meta
  // words meta(r), the padding; dst(r) is row r's record or null
  const float pad = __uint_as_float(0xFFFFFFFFu);
  auto flush = [&](uint32_t nrows, auto&& dst, auto&& meta) {
    wave_lds_sync();
    for (uint32_t w = lane; w < nrows * W; w += 64u) {
      const uint32_t r = w / W, p = w % W;
      const uint2 m = meta(r);  // on every lane (it may read other lanes' registers)
      float* d = dst(r);
      if (!d) continue;
      float4 x;
      if (4u * p + 4u <= (uint32_t)K) {
        x = *reinterpret_cast<const float4*>(st + r * RS + 4u * p);
      } else if (4u * p >= (uint32_t)K + 2u) {
        x = make_float4(pad, pad, pad, pad);
      } else {  // the word that holds the meta words (after the DOP's last two floats when K % 4 == 2)
        if (K % 4) {
          const float2 y = *reinterpret_cast<const float2*>(st + r * RS + 4u * p);
          x = make_float4(y.x, y.y, __uint_as_float(m.x), __uint_as_float(m.y));
        } else {
          x = make_float4(__uint_as_float(m.x), __uint_as_float(m.y), pad, pad);
        }
      }
      *reinterpret_cast<float4*>(d + 4u * p) = x;
    }
    wave_lds_sync();
  };
  // the second meta word of this lane's leaf (row `lane` of the leaves), loaded
  // before the DOP so that the flush does not wait for it
  const uint32_t leaf_m1 = act ? (L == 1u ? F_ROOT : 0u) | ((B.surf[tet] & 15u) ? F_SURF : 0u) | batch_body[item] : 0u;
  if (act) {
    tet_dop_at<K>(B, tet, dt, [&](uint32_t i) { return posed_vertex(B, i, iq, ip, identity != 0); }, v);
    stage_row(lane);
  }
  flush(
      64u,
      [&](uint32_t r) -> float* {  // row r: leaf r & (L-1) of chunk r >> logL
        if (tet0 + (r & ~(L - 1u)) >= nfull) return nullptr;
        return chunk_nodes + ((size_t)(r >> logL) * (2u * L) + L - 1u + (r & (L - 1u))) * S;
      },
      [&](uint32_t r) {  // its tet; its surface flag and body (lane r's leaf_m1)
        return make_uint2(F_LEAF | F_BRANCH | (tet0 + r), (uint32_t)__shfl((int)leaf_m1, (int)r, 64));
      });
  for (uint32_t t = 1; t <= logL; ++t) {
    switch (t) {
      case 1: union_level<K, 1>(v, j); break;
      case 2: union_level<K, 2>(v, j); break;
      case 3: union_level<K, 4>(v, j); break;
      case 4: union_level<K, 8>(v, j); break;
      default: union_level<K, 16>(v, j); break;
    }
    if (act && (j & ((2u << (t - 1u)) - 1u)) == 0u) stage_row((q << logL) + (1u << (logL - t)) - 1u + (j >> t));  // BFS 2^d - 1 + i
  }
  flush(
      64u,
      [&](uint32_t r) -> float* {  // row r: internal node r & (L-1) (< L-1) of chunk r >> logL
        const uint32_t i = r & (L - 1u);
        if (i == L - 1u || tet0 + (r & ~(L - 1u)) >= nfull) return nullptr;
        return chunk_nodes + ((size_t)(r >> logL) * (2u * L) + i) * S;
      },
      [&](uint32_t r) {  // its children, BFS 2i + 1 and 2i + 2 of the chunk
        const uint32_t i = r & (L - 1u), c0 = B.branch_base + ((tet0 >> logL) + (r >> logL)) * (2u * L) + 1u;
        return make_uint2(F_BRANCH | (c0 + 2u * i + 1u), (i == 0u ? F_ROOT : 0u) | (c0 + 2u * i + 2u));
      });
  if (NW == 4 && logL == 4u && B.H >= 2u) {  // L = 16: the block (one body, 256-aligned) is one super chunk of level H-2
    const uint32_t bw = threadIdx.x >> 6, btet0 = tet0 - 64u * bw;
    if (btet0 + 256u > nfull) return;  // not full (the same on every wave of the block)
    if (j == 0u) {  // every lane of a chunk holds its root; the first writes it
      float* r = roots16 + (4u * bw + q) * RS;
#pragma unroll
      for (int f = 0; f < K; ++f) r[f] = v[f];
    }
    __syncthreads();
    if (bw != 0u) return;
    super16_butterfly<K, RS>(roots16, st, lane);
    const uint32_t sc = btet0 >> 8;  // the body's super chunk of level H-2
    flush(
        15u, [&](uint32_t r) -> float* { return nodes + ((size_t)B.super_base + (size_t)sc * 16u + 1u + r) * S; },
        [&](uint32_t r) {  // its children: BFS 2r + 1 and 2r + 2 below 15 are nodes of the super chunk, the others its branch chunks' roots
          const uint32_t s0 = B.super_base + sc * 16u + 1u;
          auto child = [&](uint32_t h) { return h < 15u ? s0 + h : B.branch_base + (sc * 16u + h - 15u) * 32u + 1u; };
          return make_uint2(child(2u * r + 1u), child(2u * r + 2u));
        });
    return;
  }
  if (!fuse || tet0 + (1u << logLL) > nfull) return;  // no full super chunk in this wave (sact is false on every lane)
  if (j == 0u) stage_row(q);  // chunk roots -> rows 0 .. 64/L - 1
  wave_lds_sync();
  switch (logL) {
    case 1: super_butterfly<K, 1, RS>(st, lane, sact); break;
    case 2: super_butterfly<K, 2, RS>(st, lane, sact); break;
    default: super_butterfly<K, 3, RS>(st, lane, sact); break;
  }
  flush(
      64u >> logL,
      [&](uint32_t r) -> float* {  // row r: node r & (L-1) (< L-1) of super chunk r >> logL
        const uint32_t i = r & (L - 1u), ks = r >> logL;
        if (i == L - 1u || tet0 + ((ks + 1u) << logLL) > nfull) return nullptr;
        return nodes + ((size_t)B.super_base + (size_t)((tet0 >> logLL) + ks) * L + 1u + i) * S;
      },
      [&](uint32_t r) {  // its children: BFS 2i + 1 and 2i + 2 below L - 1 are nodes of the super chunk, the others its branch chunks' roots
        const uint32_t i = r & (L - 1u), sc = (tet0 >> logLL) + (r >> logL), s0 = B.super_base + sc * L + 1u;
        auto child = [&](uint32_t h) { return h < L - 1u ? s0 + h : B.branch_base + (sc * L + h - (L - 1u)) * (2u * L) + 1u; };
        return make_uint2(child(2u * i + 1u), child(2u * i + 2u));
      });
  if constexpr (NW == 8) {
    // L = 8 in 512-thread blocks: the block's eight waves hold the eight
    // super chunks of level H-2 (one body, 512-aligned) of one super chunk
    // of level H-3, whose roots are row 0 of each wave's stage
    const uint32_t bw = threadIdx.x >> 6, btet0 = tet0 - 64u * bw;
    if (logL != 3u || B.H < 3u || btet0 + 512u > nfull) return;  // the same on every wave of the block
    if (lane < (uint32_t)K) roots16[bw * RS + lane] = st[lane];
    __syncthreads();
    if (bw != 0u) return;
    super_butterfly<K, 3, RS>(roots16, lane, true);  // its 7 nodes -> rows 0..6
    st = roots16;
    const uint32_t sc = btet0 >> 9;  // the body's super chunk of level H-3
    flush(
        7u, [&](uint32_t r) -> float* { return nodes + ((size_t)B.super3_base + (size_t)sc * 8u + 1u + r) * S; },
        [&](uint32_t r) {  // its children: BFS 2r + 1 and 2r + 2 below 7 are its nodes, the others the roots of its level H-2 chunks
          const uint32_t s0 = B.super3_base + sc * 8u + 1u;
          auto child = [&](uint32_t h) { return h < 7u ? s0 + h : B.super_base + (sc * 8u + h - 7u) * 8u + 1u; };
          return make_uint2(child(2u * r + 1u), child(2u * r + 2u));
        });
  }
}

// A branch chunk that is not full (the last chunk of a body when L does not
// divide its tet count), by one wave: leaf DOPs from the stored world
// vertices, unions bottom-up in LDS, then one coalesced write of the chunk's
// volumes (the meta words stay).  c: the chunk.
template <int K>
struct BranchLds {
  float vol[63][K];
  uint32_t meta0[63], meta1[63];
};
template <int K>
__device__ inline void refit_partial_branch(const BodyDev& B, uint32_t c, float* __restrict__ nodes, const DirTable& dt, BranchLds<K>& sh,
                                            uint32_t lane) {
  constexpr int N = K / 2, S = rec_stride(K);
  const uint32_t L = B.L;
  const uint32_t n = (c + 1 < B.nbranch) ? L : B.nt - c * L;
  const uint32_t nn = 2 * n - 1;
  const uint32_t base = B.branch_base + c * (2 * L) + 1;  // slot of BFS node 0
  for (uint32_t i = lane; i < nn; i += 64) {
    const uint2 m = load_meta<K>(nodes, base + i);
    sh.meta0[i] = m.x; sh.meta1[i] = m.y;
  }
  wave_lds_sync();
  // leaf slabs: lanes over (node, slab)
  for (uint32_t w = lane; w < nn * N; w += 64) {
    const uint32_t i = w / N, k = w - i * N;
    const uint32_t m0 = sh.meta0[i];
    if (!(m0 & F_LEAF)) continue;
    const uint32_t tet = m0 & IDX_MASK;
    const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)tet);
    const f3 d = f3{dt.d[3 * k], dt.d[3 * k + 1], dt.d[3 * k + 2]};
    const uint32_t vi[4] = {tv.x, tv.y, tv.z, tv.w};
    float lo = FLT_MAX, hi = -FLT_MAX;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const f3 P = f3{B.X[vi[p]], B.Y[vi[p]], B.Z[vi[p]]};
      const float proj = dot3(d, P);
      lo = smin(lo, proj);
      hi = smax(hi, proj);
    }
    sh.vol[i][2 * k] = lo;
    sh.vol[i][2 * k + 1] = hi;
  }
  wave_lds_sync();
  if (n == L && (L & (L - 1)) == 0) {
    // perfect chunk: BFS order is the implicit heap, children of i at 2i+1, 2i+2
    for (int first = (int)(L / 2) - 1, cnt = (int)(L / 2); cnt >= 1; cnt >>= 1, first = (first - 1) / 2) {
      for (int w = (int)lane; w < cnt * K; w += 64) {
        const int i = first + w / K, f = w % K;
        const float a = sh.vol[2 * i + 1][f], b = sh.vol[2 * i + 2][f];
        sh.vol[i][f] = (f & 1) ? smax(a, b) : smin(a, b);
      }
      wave_lds_sync();
      if (first == 0) break;
    }
  } else {
    // partial chunk: reverse BFS sweep (children always follow their parent)
    for (int i = (int)nn - 1; i >= 0; --i) {
      const uint32_t m0 = sh.meta0[i];
      if (!(m0 & F_LEAF)) {
        const uint32_t c0 = (m0 & IDX_MASK) - base, c1 = (sh.meta1[i] & IDX_MASK) - base;
        if (lane < (uint32_t)K) {
          const float a = sh.vol[c0][lane], b = sh.vol[c1][lane];
          sh.vol[i][lane] = (lane & 1) ? smax(a, b) : smin(a, b);
        }
      }
      wave_lds_sync();
    }
  }
  for (uint32_t w = lane; w < nn * K; w += 64) {
    const uint32_t i = w / K, f = w - i * K;
    nodes[(size_t)(base + i) * S + f] = sh.vol[i][f];
  }
  wave_lds_sync();
}
// The partial chunks of the list, one wave each (bodies whose upper levels
// are refitted by launches that need the chunk's root before k_refit_top).
template <int K>
__global__ __launch_bounds__(64) void k_refit_branch(const BodyDev* __restrict__ bodies, const uint32_t* __restrict__ lbody,
                                                     const uint32_t* __restrict__ lchunk, uint32_t nlist, float* __restrict__ nodes,
                                                     DirTable dt) {
  __shared__ BranchLds<K> sh;
  const uint32_t g = blockIdx.x;
  if (g >= nlist) return;
  refit_partial_branch<K>(bodies[lbody[g]], lchunk[g], nodes, dt, sh, threadIdx.x);
}

// Super chunks (kdop_refit_tree.h:82-115): their children outside the chunk
// are resolved roots of the level below, refitted before.
//
// A full chunk (L = 2^LOGL elements) is the implicit heap over its elements
// and takes one wave, transposed: lane = g L + e holds the volume floats f =
// g + NG i (i < F) of element e, NG = 64 / L groups.  A butterfly level is
// then F = ceil(K / NG) unions per lane instead of K (4 instead of 26 at K =
// 26, L = 8), which is what a lone wave's chain of levels costs.  NG is even,
// so the floats of a lane are all lower (g even) or all upper (g odd) bounds.
// Lanes e and e ^ S hold the two children of a level (the left one without
// bit S); both end with their union, in the reference's order
// (geometry_make_union.h:9-25).
template <int K, int LOGL>
struct ChunkT {
  static constexpr uint32_t L = 1u << LOGL, NG = 64u / L, F = (K + NG - 1) / NG;
};
template <int K, int LOGL, int T>
__device__ inline void chunk_t_level(float* v, uint32_t e, uint32_t g, float* __restrict__ nodes, uint32_t base) {
  using C = ChunkT<K, LOGL>;
  constexpr uint32_t SX = 1u << (T - 1);
  const bool left = (e & SX) == 0u, upper = (g & 1u) != 0u;
#pragma unroll
  for (uint32_t i = 0; i < C::F; ++i) {
    const float y = lane_xor<SX>(v[i]);
    const float a = left ? v[i] : y, b = left ? y : v[i];
    v[i] = ubound(a, b, upper);
  }
  if ((e & ((2u << (T - 1)) - 1u)) == 0u) {  // node (2^(LOGL-T) - 1) + e / 2^T of the heap
    float* r = nodes + (size_t)(base + (1u << (LOGL - T)) - 1u + (e >> T)) * rec_stride(K);
#pragma unroll
    for (uint32_t i = 0; i < C::F; ++i) {
      const uint32_t f = g + C::NG * i;
      if (f < (uint32_t)K) r[f] = v[i];
    }
  }
}
// One full chunk whose BFS node 0 is at slot base.  src(e) gives element e's
// record {slot, row}: row ~0u reads the record from global memory, another
// row reads row `row` of the LDS table tab.  The internal nodes are stored,
// and root (an LDS row, or null) receives the chunk root.
template <int K, int LOGL, class Src>
__device__ inline void chunk_t(float* __restrict__ nodes, uint32_t base, uint32_t lane, Src&& src, const float (*tab)[K], float* root) {
  using C = ChunkT<K, LOGL>;
  const uint32_t e = lane & (C::L - 1u), g = lane >> LOGL;
  const uint2 cr = src(e);
  float v[C::F];
  if (cr.y != ~0u) {
#pragma unroll
    for (uint32_t i = 0; i < C::F; ++i) {
      const uint32_t f = g + C::NG * i;
      v[i] = f < (uint32_t)K ? tab[cr.y][f] : 0.0f;
    }
  } else {
    const float* r = nodes + (size_t)cr.x * rec_stride(K);
#pragma unroll
    for (uint32_t i = 0; i < C::F; ++i) {
      const uint32_t f = g + C::NG * i;
      v[i] = f < (uint32_t)K ? r[f] : 0.0f;
    }
  }
  if constexpr (LOGL >= 1) chunk_t_level<K, LOGL, 1>(v, e, g, nodes, base);
  if constexpr (LOGL >= 2) chunk_t_level<K, LOGL, 2>(v, e, g, nodes, base);
  if constexpr (LOGL >= 3) chunk_t_level<K, LOGL, 3>(v, e, g, nodes, base);
  if constexpr (LOGL >= 4) chunk_t_level<K, LOGL, 4>(v, e, g, nodes, base);
  if constexpr (LOGL >= 5) chunk_t_level<K, LOGL, 5>(v, e, g, nodes, base);
  if (root && e == 0u) {
#pragma unroll
    for (uint32_t i = 0; i < C::F; ++i) {
      const uint32_t f = g + C::NG * i;
      if (f < (uint32_t)K) root[f] = v[i];
    }
  }
}
// chunk_t for a chunk size known at run time (L = 2 .. 32).
template <int K, class Src>
__device__ inline void chunk_t_any(uint32_t logL, float* __restrict__ nodes, uint32_t base, uint32_t lane, Src&& src,
                                   const float (*tab)[K], float* root) {
  switch (logL) {
    case 1: chunk_t<K, 1>(nodes, base, lane, src, tab, root); break;
    case 2: chunk_t<K, 2>(nodes, base, lane, src, tab, root); break;
    case 3: chunk_t<K, 3>(nodes, base, lane, src, tab, root); break;
    case 4: chunk_t<K, 4>(nodes, base, lane, src, tab, root); break;
    default: chunk_t<K, 5>(nodes, base, lane, src, tab, root); break;
  }
}

// LDS of one wave's partial chunk: the children from outside the chunk staged
// by (internal node, side), then the internal nodes' volumes.
template <int K>
struct PartialLds {
  float ext[62][K];
  float vol[31][K];
  uint2 meta[31];
};
// The last chunk of a level when it is not full (n elements, 2 <= n < L; its
// BFS node 0 at slot base), by one wave: the reference's reverse BFS sweep
// over its n - 1 internal nodes (children always follow their parent).  The
// children from outside the chunk are first staged in LDS, a lane per
// (internal node, side), so their loads are in flight together (fetch(c, v):
// the K volume floats of record c); then lane f < K sweeps, two LDS reads a
// node.  meta(i): meta words of internal node i.  root: as for chunk_t.
template <int K, class Meta, class Fetch>
__device__ inline void partial_chunk(float* __restrict__ nodes, uint32_t base, uint32_t n, uint32_t lane, PartialLds<K>& sh,
                                     Meta&& meta, Fetch&& fetch, float* root) {
  constexpr int S = rec_stride(K);
  const uint32_t nn = n - 1;  // at most 30
  if (lane < nn) sh.meta[lane] = meta(lane);
  wave_lds_sync();
  if (lane < 2u * nn) {
    const uint32_t i = lane >> 1;
    const uint32_t c = ((lane & 1u) ? sh.meta[i].y : sh.meta[i].x) & IDX_MASK;
    if (c < base || c >= base + nn) {
      float v[K];
      fetch(c, v);
#pragma unroll
      for (int f = 0; f < K; ++f) sh.ext[lane][f] = v[f];
    }
  }
  wave_lds_sync();
  if (lane < (uint32_t)K) {
    for (int i = (int)nn - 1; i >= 0; --i) {
      const uint2 m = sh.meta[i];
      const uint32_t c0 = (m.x & IDX_MASK) - base, c1 = (m.y & IDX_MASK) - base;
      const float a = c0 < nn ? sh.vol[c0][lane] : sh.ext[2 * i][lane];
      const float b = c1 < nn ? sh.vol[c1][lane] : sh.ext[2 * i + 1][lane];
      sh.vol[i][lane] = (lane & 1) ? smax(a, b) : smin(a, b);
    }
    if (root) root[lane] = sh.vol[0][lane];
  }
  wave_lds_sync();
  for (uint32_t q = lane; q < nn * K; q += 64) {
    const uint32_t i = q / K, f = q - i * K;
    nodes[(size_t)(base + i) * S + f] = sh.vol[i][f];
  }
  wave_lds_sync();
}

// Super chunks of level h (per body h = H-1-step), chunks [cfirst, C[h]) of
// each batch item, one wave per chunk: full chunks transposed (chunk_t), the
// level's last chunk, when it is not full, by the sweep.  gpre: prefix of
// blocks per item; gitem: each block's item.
template <int K>
__global__ __launch_bounds__(64) void k_refit_super(const BodyDev* __restrict__ bodies, const uint32_t* __restrict__ batch_body,
                                                    const uint32_t* __restrict__ gpre, const uint32_t* __restrict__ gitem,
                                                    const uint32_t* __restrict__ cfirst,
                                                    const uint32_t* __restrict__ ncnk, const uint32_t* __restrict__ chunk_base,
                                                    const uint32_t* __restrict__ chunk_n, uint32_t nbatch, float* __restrict__ nodes) {
  __shared__ PartialLds<K> sh;
  const uint32_t blk = blockIdx.x;
  if (blk >= gpre[nbatch]) return;
  const uint32_t item = gitem[blk];
  const uint32_t k = blk - gpre[item];
  const BodyDev& B = bodies[batch_body[item]];
  const uint32_t L = B.L, logL = (uint32_t)(__ffs((int)L) - 1);
  const uint32_t cf = cfirst[item], nc = ncnk[item], nlast = chunk_n[item];
  const uint32_t nfc = (nc - 1u - cf) + (nlast == L ? 1u : 0u);  // full chunks from cf (cf < nc)
  const uint32_t lane = threadIdx.x;
  if (k < nfc) {
    const uint32_t base = chunk_base[item] + (cf + k) * L + 1u;  // slot of BFS node 0
    chunk_t_any<K>(logL, nodes, base, lane, [&](uint32_t e) {
      const uint2 m = load_meta<K>(nodes, base + (L >> 1) - 1u + (e >> 1));
      return make_uint2(((e & 1u) ? m.y : m.x) & IDX_MASK, ~0u);
    }, sh.vol, nullptr);
    return;
  }
  if (nlast < 2) return;
  const uint32_t base = chunk_base[item] + (nc - 1u) * L + 1u;
  partial_chunk<K>(nodes, base, nlast, lane, sh, [&](uint32_t i) { return load_meta<K>(nodes, base + i); },
                   [&](uint32_t c, float* v) {
                     NodeRec<K> r;
                     load_rec<K>(nodes, c, r);
#pragma unroll
                     for (int f = 0; f < K; ++f) v[f] = r.v[f];
                   }, nullptr);
}

// The upper levels of each batch item's tree and its root in one launch, one
// 1024-thread block per item: the super levels of few chunks, deepest first,
// then the root volume as the union of the top-level chunk roots
// (kdop_refit_tree.h:154-161).  The launch is a chain of dependent memory
// round trips, so it is built to have three: the item's entries (kTopEntries
// uint4 at ent[item * kTopEntries]: {levels, L, top roots, first top root},
// {root slot, LDS rows, meta words, partial branch chunk + 1 or 0}, then per
// level {slot of the first chunk, chunks, elements of the last chunk, first
// chunk to refit}); then, all at once, the meta words of every internal node
// of those levels' chunks and the top-root slots, into LDS; then the first
// level's children.  When the block refits every level above the body's
// partial branch chunk (the last chunk, when L does not divide the tet
// count), its last wave refits that chunk first, which saves a launch; a
// full barrier then makes its records visible to the levels.  Every
// chunk root the block computes goes to an LDS table (a row per chunk of its
// levels, chunks [cf, nc)) besides its record, and a later level or the root
// reads a child from there when the block made it, so the levels are linked
// by LDS alone: the barrier between them orders LDS only and never waits for
// the records' global stores.  A level's chunks take one wave each.
constexpr uint32_t kTopThreads = 1024;
constexpr uint32_t kTopEntries = 8;    // 2 header entries + at most 6 levels per item
constexpr uint32_t kTopRoots = 256;    // LDS rows: chunks of one item's levels
constexpr uint32_t kTopMetas = 2048;   // LDS meta words: rows x (L - 1) internal nodes
constexpr uint32_t kTopTops = 64;      // top-root slots staged in LDS (more are read from global memory)

__device__ inline void block_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS row of record `slot` if it is the root of a chunk of the block's levels
// [0, e) (lv: their entries), else ~0u (the record comes from global memory:
// a leaf, a branch chunk, or a level refitted by an earlier launch).
__device__ inline uint32_t top_row(const uint4* lv, uint32_t e, uint32_t logL, uint32_t slot) {
  uint32_t row = 0;
  for (uint32_t q = 0; q < e; ++q) {
    const uint4 l = lv[q];
    const uint32_t d = slot - (l.x + 1u), c = d >> logL;
    if (slot > l.x && (d & ((1u << logL) - 1u)) == 0u && c >= l.w && c < l.y) return row + c - l.w;
    row += l.y - l.w;
  }
  return ~0u;
}

template <int K>
struct TopLds {
  float roots[kTopRoots][K];
  uint2 meta[kTopMetas];
  uint4 ent[kTopEntries];
  uint32_t tops[kTopTops];
  union {
    PartialLds<K> part;     // the levels' partial chunks
    BranchLds<K> branch;    // the body's partial branch chunk, before the levels
  };
};

template <int K>
__global__ __launch_bounds__(kTopThreads) void k_refit_top(const uint4* __restrict__ ent, const uint32_t* __restrict__ tops,
                                                          float* __restrict__ nodes, const BodyDev* __restrict__ bodies,
                                                          const uint32_t* __restrict__ batch_body, DirTable dt) {
  constexpr int S = rec_stride(K);
  __shared__ TopLds<K> sh;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6, nw = blockDim.x >> 6;
  // round trip 1: the item's entries
  if (tid < kTopEntries) sh.ent[tid] = ent[(size_t)blockIdx.x * kTopEntries + tid];
  block_lds_barrier();
  const uint4 h0 = sh.ent[0], h1 = sh.ent[1];
  const uint32_t nlev = h0.x, L = h0.y, ntop = h0.z, top_base = h0.w, root_slot = h1.x, nmeta = h1.z;
  const uint32_t logL = (uint32_t)(__ffs((int)L) - 1), NI = L - 1u;  // NI: internal nodes per chunk
  const uint4* lv = sh.ent + 2;
  // round trip 2: meta words (row r of the levels' rows: its chunk's L - 1
  // internal nodes) and the top-root slots, every load issued before any wait
  {
    uint2 m[kTopMetas / kTopThreads];
#pragma unroll
    for (uint32_t u = 0; u < kTopMetas / kTopThreads; ++u) {
      const uint32_t g = tid + u * kTopThreads;
      m[u] = make_uint2(0u, 0u);
      if (g < nmeta) {
        uint32_t row = g / NI, e = 0, r0 = 0;
        while (row - r0 >= lv[e].y - lv[e].w) { r0 += lv[e].y - lv[e].w; ++e; }
        const uint32_t slot = lv[e].x + (lv[e].w + row - r0) * L + 1u + (g - row * NI);
        m[u] = load_meta<K>(nodes, slot);
      }
    }
    const uint32_t tv = tid < min(ntop, kTopTops) ? tops[top_base + tid] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kTopMetas / kTopThreads; ++u) {
      const uint32_t g = tid + u * kTopThreads;
      if (g < nmeta) sh.meta[g] = m[u];
    }
    if (tid < min(ntop, kTopTops)) sh.tops[tid] = tv;
  }
  block_lds_barrier();
  if (h1.w) {  // the body's partial branch chunk + 1, when this block refits every level above it
    if (w == nw - 1u) refit_partial_branch<K>(bodies[batch_body[blockIdx.x]], h1.w - 1u, nodes, dt, sh.branch, lane);
    __syncthreads();  // its records are in global memory for the levels (and its LDS free for sh.part)
  }
  uint32_t row0 = 0;  // LDS row of chunk cf of level e
  for (uint32_t e = 0; e < nlev; ++e) {
    const uint4 l = lv[e];  // cbase, nc, nlast, cf
    const uint32_t nfc = (l.y - 1u - l.w) + (l.z == L ? 1u : 0u);
    const uint32_t ntask = nfc + (l.z < L && l.z >= 2u ? 1u : 0u);
    for (uint32_t k = w; k < ntask; k += nw) {
      if (k < nfc) {
        const uint32_t base = l.x + (l.w + k) * L + 1u, mrow = (row0 + k) * NI;
        chunk_t_any<K>(logL, nodes, base, lane, [&](uint32_t el) {
          const uint2 m = sh.meta[mrow + (L >> 1) - 1u + (el >> 1)];
          const uint32_t c = ((el & 1u) ? m.y : m.x) & IDX_MASK;
          return make_uint2(c, top_row(lv, e, logL, c));
        }, sh.roots, sh.roots[row0 + k]);
      } else {
        const uint32_t base = l.x + (l.y - 1u) * L + 1u, prow = row0 + (l.y - 1u - l.w);
        partial_chunk<K>(nodes, base, l.z, lane, sh.part, [&](uint32_t i) { return sh.meta[prow * NI + i]; },
                         [&](uint32_t c, float* v) {
                           const uint32_t row = top_row(lv, e, logL, c);
                           if (row != ~0u) {
#pragma unroll
                             for (int f = 0; f < K; ++f) v[f] = sh.roots[row][f];
                           } else {
                             NodeRec<K> r;
                             load_rec<K>(nodes, c, r);
#pragma unroll
                             for (int f = 0; f < K; ++f) v[f] = r.v[f];
                           }
                         }, sh.roots[prow]);
      }
    }
    block_lds_barrier();
    row0 += l.y - l.w;
  }
  if (w != 0 || lane >= (uint32_t)K) return;
  float acc = 0.0f;
  for (uint32_t t = 0; t < ntop; ++t) {
    const uint32_t c = t < kTopTops ? sh.tops[t] : tops[top_base + t], row = top_row(lv, nlev, logL, c);
    const float x = row != ~0u ? sh.roots[row][lane] : nodes[(size_t)c * S + lane];
    acc = t == 0 ? x : ((lane & 1) ? smax(acc, x) : smin(acc, x));
  }
  nodes[(size_t)root_slot * S + lane] = acc;
}

// Bodies with chunks wider than 32 leaves (a per-body mem_bytes above about
// 8 kB, or one chunk of every tet): the topology comes from the host, and the
// refit runs from two host-built lists.  Leaves: one thread per tet writes
// its DOP at leaf_slot[tet].
template <int K>
__global__ __launch_bounds__(256) void k_refit_tet_slots(const BodyDev* __restrict__ bodies, uint32_t body,
                                                         const uint32_t* __restrict__ leaf_slot, float* __restrict__ nodes,
                                                         DirTable dt) {
  const BodyDev& B = bodies[body];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.nt) return;
  float v[K];
  tet_dop<K>(B, t, dt, v);
  store_vol<K>(nodes, leaf_slot[t], v);
}
// Internal nodes of one height (1 + the larger height of the two children,
// leaves 0), one thread each: the union of its children (left, right), in
// the reference's order (geometry_make_union.h:9-25).  Launched height by
// height, so every child is final before its parent reads it.
template <int K>
__global__ __launch_bounds__(256) void k_refit_union_list(const uint32_t* __restrict__ list, uint32_t n, float* __restrict__ nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t slot = list[i];
  const uint2 m = load_meta<K>(nodes, slot);
  NodeRec<K> a, b;
  load_rec<K>(nodes, m.x & IDX_MASK, a);
  load_rec<K>(nodes, m.y & IDX_MASK, b);
  float v[K];
#pragma unroll
  for (int f = 0; f < K; ++f) v[f] = (f & 1) ? smax(a.v[f], b.v[f]) : smin(a.v[f], b.v[f]);
  store_vol<K>(nodes, slot, v);
}
// Meta words of a host-built topology: record i of the body gets m0, m1 =
// meta[2 i], meta[2 i + 1] (the volumes are written by the refit).
__global__ void k_write_meta(const uint32_t* __restrict__ meta, uint32_t nnodes, uint32_t node_base, int K, float* __restrict__ nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  uint32_t* m = reinterpret_cast<uint32_t*>(nodes + (size_t)(node_base + i) * rec_stride(K) + K);
  m[0] = meta[2 * (size_t)i];
  m[1] = meta[2 * (size_t)i + 1];
}

// ============================================================== traversal
struct TravArgs {
  const float* nodes;
  const BodyDev* bodies;
  const uint2* pairs;        // body pair ids
  const uint32_t* tops;
  CounterStripe* stripes;    // kStripes striped counters
  unsigned int* overflow;    // bit 0 frontier, 1 P1, 2 P2, 3 contacts
  LeafPair* p1;              // may be null (P1 not kept)
  LeafPair* p2;
  uint32_t* p1_fill;
  uint32_t* p2_fill;
  uint32_t p1_cap, p2_cap;
  uint32_t npairs;
  uint8_t* touched;          // record-counting calls only: byte per node slot, set when a round reads the record
};

// Which sides of an overlapping node pair (x, y) expand: the reference's rule
// inside branch chunks (kdop_tandem_traversal.h:122-154: both internal -> 4
// pairs, one leaf -> the other side's 2 children), with branch roots treated
// as leaves while the other side is still above the branch level
// (super-chunk levels, which the reference does not traverse).
__device__ inline void expand_rule(uint32_t mx0, uint32_t my0, bool& ex, bool& ey) {
  const bool xb = mx0 & F_BRANCH, yb = my0 & F_BRANCH;
  if (!xb && !yb) { ex = true; ey = true; }
  else if (!xb) { ex = true; ey = false; }
  else if (!yb) { ex = false; ey = true; }
  else { ex = !(mx0 & F_LEAF); ey = !(my0 & F_LEAF); }
}
// Work item for the overlapping pair (x at slot xi, y at slot yi).
template <int K>
__device__ inline Item make_item(const NodeRec<K>& X, uint32_t xi, const NodeRec<K>& Y, uint32_t yi, uint32_t pair);
// The same from the two records' meta words (m0, m1).
__device__ inline Item make_item_meta(uint2 mx, uint32_t xi, uint2 my, uint32_t yi, uint32_t pair) {
  bool ex, ey;
  expand_rule(mx.x, my.x, ex, ey);
  Item it;
  it.a0 = ex ? (I_EXP | (mx.x & IDX_MASK)) : xi;
  it.a1 = ex ? (mx.y & IDX_MASK) : 0u;
  it.b0 = ey ? (I_EXP | (my.x & IDX_MASK)) : yi;
  it.b1 = ey ? (my.y & IDX_MASK) : 0u;
  it.pair = pair;
  return it;
}
template <int K>
__device__ inline Item make_item(const NodeRec<K>& X, uint32_t xi, const NodeRec<K>& Y, uint32_t yi, uint32_t pair) {
  return make_item_meta(make_uint2(X.m0, X.m1), xi, make_uint2(Y.m0, Y.m1), yi, pair);
}

// Result of expanding one work item: new items and three 4-bit masks over the
// child combinations k = 2 i + j (child i of side a, child j of side b).
struct Expansion {
  Item kids[4];           // kids[k] valid where bit k of push is set
  uint32_t ta[2], tb[2];  // tet ids when the child is a tet leaf
  uint32_t push;          // overlapping combinations that become new items
  uint32_t leaf;          // overlapping tet-leaf pairs (P1)
  uint32_t surf;          // subset of leaf where both tets have a surface face (P2)
  uint32_t tests, descent, broot;
};

// Test the child pairs of one item (every test is one node-pair test of the
// BVTT; descent counts those below branch-root pairs, broot the overlapping
// branch-root pairs).
template <int K>
__device__ inline void expand_item(const float* __restrict__ nodes, const Item& it, Expansion& x) {
  const bool exp_a = it.a0 & I_EXP, exp_b = it.b0 & I_EXP;
  const uint32_t ca0 = it.a0 & IDX_MASK, ca1 = it.a1, cb0 = it.b0 & IDX_MASK, cb1 = it.b1;
  NodeRec<K> A0, A1, B0, B1;
  A1.m0 = A1.m1 = B1.m0 = B1.m1 = 0u;
  load_rec<K>(nodes, ca0, A0);
  if (exp_a) load_rec<K>(nodes, ca1, A1);
  load_rec<K>(nodes, cb0, B0);
  if (exp_b) load_rec<K>(nodes, cb1, B1);
  x.ta[0] = A0.m0 & IDX_MASK; x.ta[1] = A1.m0 & IDX_MASK;
  x.tb[0] = B0.m0 & IDX_MASK; x.tb[1] = B1.m0 & IDX_MASK;
  x.push = 0; x.leaf = 0; x.surf = 0;
  x.tests = 0; x.descent = 0; x.broot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = k >> 1, j = k & 1;
    if ((i == 1 && !exp_a) || (j == 1 && !exp_b)) continue;
    const NodeRec<K>& X = i ? A1 : A0;
    const NodeRec<K>& Y = j ? B1 : B0;
    ++x.tests;
    const bool xb = X.m0 & F_BRANCH, yb = Y.m0 & F_BRANCH;
    const bool both_roots = xb && yb && (X.m1 & F_ROOT) && (Y.m1 & F_ROOT);
    if (xb && yb && !both_roots) ++x.descent;
    if (!dop_overlap<K>(X.v, Y.v)) continue;
    if (both_roots) ++x.broot;
    if ((X.m0 & F_LEAF) && (Y.m0 & F_LEAF)) {
      x.leaf |= 1u << k;
      if ((X.m1 & F_SURF) && (Y.m1 & F_SURF)) x.surf |= 1u << k;
    } else {
      x.push |= 1u << k;
      x.kids[k] = make_item<K>(X, i ? ca1 : ca0, Y, j ? cb1 : cb0, it.pair);
    }
  }
}

// ------------------------------------------- cooperative expansion (4 lanes)
// A work item expanded by the 4 lanes of a quad: lane r of the quad loads the
// 16-byte words r and r + 4 of every record it needs (so one load instruction
// covers 16 records in whole 64-byte runs instead of 64 scattered records),
// tests the slabs held in those words for all 4 child combinations, and the
// quad ORs its separation bits; lane r then owns combination k = r (child
// i = r >> 1 of side a, child j = r & 1 of side b): its test, its counters and
// the one item or leaf pair it may produce.  The meta words come from the
// lane that loaded them by a quad broadcast.
template <int K>
struct Quad {
  static constexpr int QU = (K + 2 + 3) / 4;  // 16-byte words holding vol[K], m0, m1
  static constexpr int NQ = (QU + 3) / 4;     // of them per lane
  static constexpr int MQ = K / 4;            // the word holding m0, m1
  static constexpr int ML = MQ & 3;           // its lane in the quad
  static constexpr int MT = MQ >> 2;          // and that lane's load slot
  static constexpr bool MZ = (K % 4) != 0;    // m0, m1 in .z .w (else .x .y)
};

template <int CTRL>
__device__ inline uint32_t quad_dpp(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false); }

// The quad's view of one record: this lane's words (slot t: word r + 4 t).
template <int K>
struct QuadRec {
  float4 w[Quad<K>::NQ];
};
// A lane loads only its slots that hold part of the record (lanes past the
// used words keep what they had); lane_base: the pool as 16-byte words,
// offset by this lane's first word.
// Every lane loads all its slots, so the loads take no branches (measured
// faster than predicating the lanes and records that are not needed): a word
// past the used ones is the record's padding, or for K = 6 its last word
// again; quad_sep masks the words that hold no slab.  lane_base: the pool as
// 16-byte words, offset by this lane's first word.
template <int K>
__device__ inline void load_quad(const float4* __restrict__ lane_base, uint32_t idx, QuadRec<K>& q) {
  constexpr uint32_t RW = rec_stride(K) / 4;  // 16-byte words per record
  const float4* p = lane_base + (size_t)idx * RW;
#pragma unroll
  for (int t = 0; t < Quad<K>::NQ; ++t) q.w[t] = p[4 * t];
}
// This lane's first word of a record (lanes past the record's words for K = 6
// read its last word; they test no slab).
template <int K>
__device__ inline const float4* quad_lane_base(const float* __restrict__ nodes, uint32_t r) {
  constexpr uint32_t RW = rec_stride(K) / 4;
  return reinterpret_cast<const float4*>(nodes) + (r < RW ? r : RW - 1u);
}
// This lane's share of the strict separation test of X and Y: the slabs held
// in its words (lo0 hi0 lo1 hi1 per word; lanes past the volume test none).
template <int K>
__device__ inline bool quad_sep(const QuadRec<K>& X, const QuadRec<K>& Y, uint32_t r) {
  bool s = false;
#pragma unroll
  for (int t = 0; t < Quad<K>::NQ; ++t) {
    const uint32_t q = r + 4u * t;
    const float4 x = X.w[t], y = Y.w[t];
    const bool v0 = 4u * q + 1u < (uint32_t)K, v1 = 4u * q + 3u < (uint32_t)K;
    s |= v0 & ((x.y < y.x) | (y.y < x.x));
    s |= v1 & ((x.w < y.z) | (y.w < x.z));
  }
  return s;
}
template <int K>
__device__ inline uint2 quad_meta(const QuadRec<K>& X) {
  const float4 m = X.w[Quad<K>::MT];
  const uint32_t m0 = __float_as_uint(Quad<K>::MZ ? m.z : m.x), m1 = __float_as_uint(Quad<K>::MZ ? m.w : m.y);
  constexpr int B = Quad<K>::ML * 0x55;  // quad_perm [ML, ML, ML, ML]
  return make_uint2(quad_dpp<B>(m0), quad_dpp<B>(m1));
}

// What lane r of a quad makes of its child combination.
struct QuadOut {
  Item kid;             // valid when push
  uint32_t ta, tb;      // tet ids when leaf
  bool test, descent, broot, leaf, surf, push;
};

// Every lane of the wave must call it (quad-wide DPP); live: the quad holds
// an item (quad-uniform; a dead quad's item is all zeros, so its loads read
// record 0).  A side that does not expand loads its one record twice, which
// keeps the loads free of branches.  A0 .. B1: the caller's registers for the
// four records.
// Record slot s (0: A's first, 1: A's second, 2: B's first, 3: B's second
// node) that expand_quad loads for an item: a side that does not expand
// names its one node twice.
__device__ inline uint32_t quad_record(const Item& it, uint32_t s) {
  const bool exp_a = (it.a0 & I_EXP) != 0u, exp_b = (it.b0 & I_EXP) != 0u;
  const uint32_t ca0 = it.a0 & IDX_MASK, cb0 = it.b0 & IDX_MASK;
  const uint32_t x0 = s < 2u ? ca0 : cb0, x1 = s < 2u ? (exp_a ? it.a1 : ca0) : (exp_b ? it.b1 : cb0);
  return (s & 1u) ? x1 : x0;
}
template <int K>
__device__ inline void expand_quad_test(const Item& it, bool live, const QuadRec<K>& A0, const QuadRec<K>& A1, const QuadRec<K>& B0,
                                        const QuadRec<K>& B1, QuadOut& o);
template <int K>
__device__ inline void expand_quad(const float4* __restrict__ lane_base, const Item& it, bool live, QuadRec<K>& A0, QuadRec<K>& A1,
                                   QuadRec<K>& B0, QuadRec<K>& B1, QuadOut& o) {
  const bool exp_a = (it.a0 & I_EXP) != 0u, exp_b = (it.b0 & I_EXP) != 0u;
  const uint32_t ca0 = it.a0 & IDX_MASK, cb0 = it.b0 & IDX_MASK;
  const uint32_t ca1 = exp_a ? it.a1 : ca0, cb1 = exp_b ? it.b1 : cb0;
  load_quad<K>(lane_base, ca0, A0);
  load_quad<K>(lane_base, ca1, A1);
  load_quad<K>(lane_base, cb0, B0);
  load_quad<K>(lane_base, cb1, B1);
  expand_quad_test<K>(it, live, A0, A1, B0, B1, o);
}
// The tests of expand_quad on records already in this lane's words.
template <int K>
__device__ inline void expand_quad_test(const Item& it, bool live, const QuadRec<K>& A0, const QuadRec<K>& A1, const QuadRec<K>& B0,
                                        const QuadRec<K>& B1, QuadOut& o) {
  const uint32_t r = (uint32_t)lane_id() & 3u, i = r >> 1, j = r & 1u;
  const bool exp_a = (it.a0 & I_EXP) != 0u, exp_b = (it.b0 & I_EXP) != 0u;
  const uint32_t ca0 = it.a0 & IDX_MASK, cb0 = it.b0 & IDX_MASK;
  const uint32_t ca1 = exp_a ? it.a1 : ca0, cb1 = exp_b ? it.b1 : cb0;
  // separation bit of combination k in bit k, ORed over the quad
  uint32_t sep = (quad_sep<K>(A0, B0, r) ? 1u : 0u) | (quad_sep<K>(A0, B1, r) ? 2u : 0u) | (quad_sep<K>(A1, B0, r) ? 4u : 0u) |
                 (quad_sep<K>(A1, B1, r) ? 8u : 0u);
  sep |= quad_dpp<0xB1>(sep);  // quad_perm [1, 0, 3, 2]
  sep |= quad_dpp<0x4E>(sep);  // quad_perm [2, 3, 0, 1]
  const uint2 ma0 = quad_meta<K>(A0), ma1 = quad_meta<K>(A1), mb0 = quad_meta<K>(B0), mb1 = quad_meta<K>(B1);
  const uint2 X = i ? ma1 : ma0, Y = j ? mb1 : mb0;
  const uint32_t xi = i ? ca1 : ca0, yi = j ? cb1 : cb0;
  o.test = live && (i == 0u || exp_a) && (j == 0u || exp_b);
  const bool xb = (X.x & F_BRANCH) != 0u, yb = (Y.x & F_BRANCH) != 0u;
  const bool both_roots = xb && yb && (X.y & F_ROOT) && (Y.y & F_ROOT);
  o.descent = o.test && xb && yb && !both_roots;
  const bool ov = o.test && ((sep >> r) & 1u) == 0u;
  o.broot = ov && both_roots;
  o.leaf = ov && (X.x & F_LEAF) && (Y.x & F_LEAF);
  o.surf = o.leaf && (X.y & F_SURF) && (Y.y & F_SURF);
  o.push = ov && !o.leaf;
  o.ta = X.x & IDX_MASK;
  o.tb = Y.x & IDX_MASK;
  bool ex, ey;
  expand_rule(X.x, Y.x, ex, ey);
  o.kid.a0 = ex ? (I_EXP | (X.x & IDX_MASK)) : xi;
  o.kid.a1 = ex ? (X.y & IDX_MASK) : 0u;
  o.kid.b0 = ey ? (I_EXP | (Y.x & IDX_MASK)) : yi;
  o.kid.b1 = ey ? (Y.y & IDX_MASK) : 0u;
  o.kid.pair = it.pair;
}

// Gather roof (measurement only): quads load records idx[j] of the node pool
// exactly as expand_quad loads an item's records (4 records per quad per
// iteration, each lane its words r and r + 4 of the used ones), and fold them
// into a checksum so no load is dropped.  The rate of record bytes this
// reaches over random indices bounds what the traversal's loads can reach.
template <int K>
__global__ __launch_bounds__(256) void k_gather_probe(const float* __restrict__ nodes, const uint32_t* __restrict__ idx, uint32_t nidx,
                                                      uint32_t* __restrict__ out) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, r = tid & 3u;
  const uint32_t quad = tid >> 2, nquads = (gridDim.x * blockDim.x) >> 2;
  const float4* __restrict__ base = quad_lane_base<K>(nodes, r);
  QuadRec<K> q[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int t = 0; t < Quad<K>::NQ; ++t) q[u].w[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t acc = 0;
  for (uint32_t j = 4u * quad; j + 3u < nidx; j += 4u * nquads) {
#pragma unroll
    for (int u = 0; u < 4; ++u) load_quad<K>(base, idx[j + u], q[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < Quad<K>::NQ; ++t)
        acc ^= __float_as_uint(q[u].w[t].x) ^ __float_as_uint(q[u].w[t].y) ^ __float_as_uint(q[u].w[t].z) ^ __float_as_uint(q[u].w[t].w);
  }
  out[tid] = acc;
}

// Write the leaf pairs of one lane (P1 when kept, P2 for the surface subset).
// Every lane of the wave must call it.
__device__ inline void emit_leaves(const TravArgs& T, uint32_t pair, const Expansion& x) {
  if (T.p1) {
    const uint32_t o1 = wave_reserve(T.p1_fill, __popc(x.leaf));
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if ((x.leaf >> k) & 1u) {
        const uint32_t o = o1 + __popc(x.leaf & ((1u << k) - 1u));
        if (o < T.p1_cap) T.p1[o] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
        else atomicOr(T.overflow, 2u);
      }
  }
  const uint32_t o2 = wave_reserve(T.p2_fill, __popc(x.surf));
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((x.surf >> k) & 1u) {
      const uint32_t o = o2 + __popc(x.surf & ((1u << k) - 1u));
      if (o < T.p2_cap) T.p2[o] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
      else atomicOr(T.overflow, 4u);
    }
}

// P1 part of emit_leaves (every lane of the wave calls it; nothing when P1 is
// not kept).
__device__ inline void emit_p1(const TravArgs& T, uint32_t pair, const Expansion& x) {
  if (!T.p1) return;
  const uint32_t o1 = wave_reserve(T.p1_fill, __popc(x.leaf));
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((x.leaf >> k) & 1u) {
      const uint32_t o = o1 + __popc(x.leaf & ((1u << k) - 1u));
      if (o < T.p1_cap) T.p1[o] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
      else atomicOr(T.overflow, 2u);
    }
}

// Surface leaf pairs (P2) staged per block in LDS: a wave-wide reservation on
// the one P2 counter for every wave that found some serializes across the
// XCDs at the leaf levels of a wide traversal; staged, a block reserves once
// per kP2Stage - 1024 pairs (an iteration of 256 lanes adds at most 1024).
constexpr uint32_t kP2Stage = 2048;
struct P2Stage {
  LeafPair p[kP2Stage];
  uint32_t n, base;
};
__device__ inline void stage_p2(P2Stage& S, uint32_t pair, const Expansion& x) {
  const uint32_t ns = __popc(x.surf);
  if (!ns) return;
  uint32_t pos = atomicAdd(&S.n, ns);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((x.surf >> k) & 1u) S.p[pos++] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
}
// One lane's surface pair (expand_quad: at most one per lane), with one LDS
// reservation per wave.  Every lane of the wave must call it.
__device__ inline void stage_p2_lane(P2Stage& S, uint32_t pair, bool surf, uint32_t ta, uint32_t tb) {
  const unsigned long long m = __ballot(surf);
  if (m == 0ull) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(&S.n, (uint32_t)__popcll(m));
  base = lane0(base);
  if (surf) S.p[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = LeafPair{pair, ta, tb};
}
// Write the stage to P2 when forced (the block's end) or when the next
// iteration might not fit.  Every thread of the block calls it.
__device__ inline void flush_p2(const TravArgs& T, P2Stage& S, bool force) {
  __syncthreads();
  const uint32_t m = S.n;
  if (m == 0u || (!force && m <= kP2Stage - 1024u)) return;
  if (threadIdx.x == 0) {
    S.base = atomicAdd(T.p2_fill, m);
    if (S.base + m > T.p2_cap) atomicOr(T.overflow, 4u);
  }
  __syncthreads();
  const uint32_t b = S.base;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
    if (b + i < T.p2_cap) T.p2[b + i] = S.p[i];
  __syncthreads();
  if (threadIdx.x == 0) S.n = 0u;
  __syncthreads();
}

// Write this lane's new items at slot o of `out` (one wave-wide reservation).
__device__ inline void emit_items(const Expansion& x, Item* __restrict__ out, uint32_t o, uint32_t cap, unsigned int* overflow) {
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((x.push >> k) & 1u) {
      const uint32_t slot = o + __popc(x.push & ((1u << k) - 1u));
      if (slot < cap) out[slot] = x.kids[k];
      else atomicOr(overflow, 1u);
    }
}

// Seed descriptor of one body pair: every chunk root of level hA of body A
// (tops[a_off .. a_off+na)) against every chunk root of level hB of body B.
// One body pair's seeds: chunk roots [a_off, a_off + na) x [b_off, b_off +
// nb) of the context's root table, seed indices from base; with the pair's
// root records and branch-chunk counts, so that a seed thread needs no body
// descriptor before its first record loads.
struct SeedDesc {
  uint32_t a_off, na, b_off, nb, base;
  uint32_t root_a, root_b;    // root records of body A and body B
  uint32_t nbr_a, nbr_b;      // branch chunks C_A, C_B (the reference's C_A x C_B branch-pair loop)
};

// Seed: one thread per (chunk root of A, chunk root of B) pair of the chosen
// levels, for every body pair whose root volumes overlap (the root-vs-root
// test is kdop_tandem_traversal.h:164-165).  Seeding below the top level
// only skips tests among super-chunk nodes, which the reference never
// makes; the pairs found and every test at or below the branch roots are
// unchanged.  EXPAND: a seed pair that overlaps is expanded at once, as the
// first breadth-first level would expand it (its child pairs tested, the
// overlapping ones written as items of level 1), which saves that level's
// launch; the tests made are the same.
template <int K, bool EXPAND>
__global__ __launch_bounds__(256) void k_seed(TravArgs T, const SeedDesc* __restrict__ sd, uint32_t total, Item* __restrict__ out,
                                             uint32_t* __restrict__ out_fill, uint32_t out_cap) {
  __shared__ uint32_t red[5];
  unsigned long long tests = 0, broot = 0, p1s = 0, roots = 0, bpairs = 0, descent = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t g = base + threadIdx.x;
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    uint32_t pair = 0;
    if (g < total) {
      uint32_t lo = 0, hi = T.npairs;
      while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (sd[m].base <= g) lo = m; else hi = m;
      }
      pair = lo;
      const SeedDesc d = sd[pair];
      const uint32_t local = g - d.base;
      // the root records and the seed's chunk roots load together; the seed
      // counts only when the roots overlap
      const uint32_t i = local / d.nb, j = local - i * d.nb;
      const uint32_t xa = T.tops[d.a_off + i], yb = T.tops[d.b_off + j];
      NodeRec<K> ra, rb, X, Y;
      load_rec<K>(T.nodes, d.root_a, ra);
      load_rec<K>(T.nodes, d.root_b, rb);
      load_rec<K>(T.nodes, xa, X);
      load_rec<K>(T.nodes, yb, Y);
      const bool ov = dop_overlap<K>(ra.v, rb.v);
      if (local == 0) {
        ++roots;
        if (ov) bpairs += (unsigned long long)d.nbr_a * d.nbr_b;
      }
      if (ov) {
        x.ta[0] = X.m0 & IDX_MASK;
        x.tb[0] = Y.m0 & IDX_MASK;
        ++tests;
        const bool both_roots = (X.m0 & F_BRANCH) && (Y.m0 & F_BRANCH) && (X.m1 & F_ROOT) && (Y.m1 & F_ROOT);
        if (dop_overlap<K>(X.v, Y.v)) {
          if (both_roots) ++broot;
          if ((X.m0 & F_LEAF) && (Y.m0 & F_LEAF)) {
            x.leaf = 1u;
            x.surf = ((X.m1 & F_SURF) && (Y.m1 & F_SURF)) ? 1u : 0u;
          } else {
            x.push = 1u;
            x.kids[0] = make_item<K>(X, xa, Y, yb, pair);
          }
        }
      }
    }
    p1s += x.leaf;
    emit_leaves(T, pair, x);
    if (EXPAND) {
      Expansion x2;
      x2.push = x2.leaf = x2.surf = 0;
      x2.tests = x2.descent = x2.broot = 0;
      if (x.push) expand_item<K>(T.nodes, x.kids[0], x2);
      tests += x2.tests; descent += x2.descent; broot += x2.broot; p1s += __popc(x2.leaf);
      const uint32_t o = block_reserve(out_fill, __popc(x2.push), red);
      emit_items(x2, out, o, out_cap, T.overflow);
      emit_leaves(T, pair, x2);
    } else {
      const uint32_t o = block_reserve(out_fill, x.push, red);
      emit_items(x, out, o, out_cap, T.overflow);
    }
  }
  CounterStripe* st = stripe_of(T.stripes, wave_global());
  wave_add64(&st->root_tests, roots);
  wave_add64(&st->tests, tests + roots);
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_pairs, bpairs);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
}

// k_seed<K, true> with four lanes per seed pair (expand_quad): each quad
// loads the root records and the seed's chunk roots as 16-byte words (8 loads
// per lane instead of 32), tests both pairs, and expands an overlapping seed
// pair itself.  The quad's first lane counts and writes what belongs to the
// seed pair; every lane its own child combination.  The same tests, counts,
// pairs and items as k_seed<K, true>; the persistent strategy's seeding.
template <int K>
__global__ __launch_bounds__(256) void k_seed_quad(TravArgs T, const SeedDesc* __restrict__ sd, uint32_t total, Item* __restrict__ out,
                                                  uint32_t* __restrict__ out_fill, uint32_t out_cap) {
  __shared__ uint32_t red[5];
  unsigned long long tests = 0, broot = 0, p1s = 0, roots = 0, bpairs = 0, descent = 0;
  const uint32_t r = threadIdx.x & 3u;
  const bool lead = r == 0u;
  const float4* __restrict__ qbase = quad_lane_base<K>(T.nodes, r);
  const uint32_t spb = blockDim.x / 4u;  // seed pairs per block pass
  for (uint32_t base = blockIdx.x * spb; base < total; base += gridDim.x * spb) {
    const uint32_t g = base + (threadIdx.x >> 2);
    const bool live = g < total;
    uint32_t pair = 0, xa = 0, yb = 0, ra = 0, rb = 0, local = 0;
    unsigned long long nbr = 0;
    if (live) {
      uint32_t lo = 0, hi = T.npairs;
      while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (sd[m].base <= g) lo = m; else hi = m;
      }
      pair = lo;
      const SeedDesc d = sd[pair];
      local = g - d.base;
      const uint32_t i = local / d.nb, j = local - i * d.nb;
      xa = T.tops[d.a_off + i];
      yb = T.tops[d.b_off + j];
      ra = d.root_a;
      rb = d.root_b;
      nbr = (unsigned long long)d.nbr_a * d.nbr_b;
    }
    // the root records and the seed's chunk roots load together (a dead quad reads record 0)
    QuadRec<K> RA, RB, X, Y;
    load_quad<K>(qbase, ra, RA);
    load_quad<K>(qbase, rb, RB);
    load_quad<K>(qbase, xa, X);
    load_quad<K>(qbase, yb, Y);
    uint32_t sep = (quad_sep<K>(RA, RB, r) ? 1u : 0u) | (quad_sep<K>(X, Y, r) ? 2u : 0u);
    sep |= quad_dpp<0xB1>(sep);  // quad_perm [1, 0, 3, 2]
    sep |= quad_dpp<0x4E>(sep);  // quad_perm [2, 3, 0, 1]
    const uint2 mx = quad_meta<K>(X), my = quad_meta<K>(Y);
    const bool ov = live && (sep & 1u) == 0u;  // the seed counts only when the roots overlap
    if (lead && live && local == 0u) {
      ++roots;
      if (ov) bpairs += nbr;
    }
    bool push = false, leaf = false, surf = false;
    Item kid{0, 0, 0, 0, 0};
    if (ov) {
      if (lead) ++tests;
      const bool both_roots = (mx.x & F_BRANCH) && (my.x & F_BRANCH) && (mx.y & F_ROOT) && (my.y & F_ROOT);
      if ((sep & 2u) == 0u) {
        if (lead && both_roots) ++broot;
        if ((mx.x & F_LEAF) && (my.x & F_LEAF)) {
          leaf = true;
          surf = (mx.y & F_SURF) && (my.y & F_SURF);
        } else {
          push = true;
          kid = make_item_meta(mx, xa, my, yb, pair);
        }
      }
    }
    // a seed pair of two tet leaves (one-tet bodies): written by the quad's first lane
    const bool sl = lead && leaf, ss = lead && surf;
    p1s += sl ? 1u : 0u;
    if (T.p1) {
      const uint32_t o1 = wave_reserve(T.p1_fill, sl ? 1u : 0u);
      if (sl) {
        if (o1 < T.p1_cap) T.p1[o1] = LeafPair{pair, mx.x & IDX_MASK, my.x & IDX_MASK};
        else atomicOr(T.overflow, 2u);
      }
    }
    {
      const uint32_t o2 = wave_reserve(T.p2_fill, ss ? 1u : 0u);
      if (ss) {
        if (o2 < T.p2_cap) T.p2[o2] = LeafPair{pair, mx.x & IDX_MASK, my.x & IDX_MASK};
        else atomicOr(T.overflow, 4u);
      }
    }
    // the overlapping seed pair's children (quad-uniform push)
    QuadRec<K> A0, A1, B0, B1;
    QuadOut o;
    expand_quad<K>(qbase, kid, push, A0, A1, B0, B1, o);
    tests += o.test; descent += o.descent; broot += o.broot; p1s += o.leaf;
    const uint32_t slot = block_reserve(out_fill, o.push ? 1u : 0u, red);
    if (o.push) {
      if (slot < out_cap) out[slot] = o.kid;
      else atomicOr(T.overflow, 1u);
    }
    if (T.p1) {
      const uint32_t o1 = wave_reserve(T.p1_fill, o.leaf ? 1u : 0u);
      if (o.leaf) {
        if (o1 < T.p1_cap) T.p1[o1] = LeafPair{pair, o.ta, o.tb};
        else atomicOr(T.overflow, 2u);
      }
    }
    const uint32_t o2 = wave_reserve(T.p2_fill, o.surf ? 1u : 0u);
    if (o.surf) {
      if (o2 < T.p2_cap) T.p2[o2] = LeafPair{pair, o.ta, o.tb};
      else atomicOr(T.overflow, 4u);
    }
  }
  CounterStripe* st = stripe_of(T.stripes, wave_global());
  wave_add64(&st->root_tests, roots);
  wave_add64(&st->tests, tests + roots);
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_pairs, bpairs);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
}

// ============================================================== narrow phase
struct NarrowArgs {
  const BodyDev* bodies;
  const uint2* pairs;
  const LeafPair* p2;
  const uint32_t* p2_fill;
  uint32_t p2_cap;
  kdop_contact* contacts;
  uint32_t* c_fill;
  uint32_t c_cap;
  NarrowRec* recs;
  uint32_t* r_fill;
  uint32_t r_cap;
  CounterStripe* stripes;
  unsigned int* overflow;
  // gate list of the two-pass narrow phase (P2 index | face << 30) and its fill
  uint32_t* gate;
  uint32_t* g_fill;
  // when non-null: runs only if every word of pending[0..kSegs) is zero (the
  // persistent traversal left no spill behind; else the host drains it and
  // launches the narrow phase again)
  const uint32_t* pending;
  uint32_t pending_stride;  // words between consecutive pending entries
  // when non-null: runs only if *done is nonzero (the level-synchronous
  // traversal reached an empty frontier)
  const uint32_t* done;
};

// World vertex v of B (the refit's (x, y, z, 0) copy: one 16-byte load).
__device__ inline f3 world_vertex(const BodyDev& B, uint32_t v) {
  const float4 w = B.W[v];
  return f3{w.x, w.y, w.z};
}
__device__ inline void gather_tet(const BodyDev& B, uint32_t t, f3* P) {
  const uint4 tv = *reinterpret_cast<const uint4*>(B.tets + 4 * (size_t)t);
  const uint32_t v[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) P[q] = world_vertex(B, v[q]);
}

// SAT gate (overlap/geometry_overlap_tetrahedron_tetrahedron.h:50-290) and
// the most opposing face pair (contacts ...tetrahedron.h:464-533) of surface
// pair i with the 8 lanes of its group: each lane takes 5-6 of the 44 SAT axes
// and 2 of the 16 face pairs; the serial loops' results (any separating axis;
// the first face pair, in a-major order, with the smallest dot product) come
// from ballots over the group.  V, U: the group's LDS for the two tets'
// vertices and face planes.  Returns the A face, or -1 (same in all 8 lanes).
__device__ inline int gate_group(const NarrowArgs& A, f3* V, UPlane* U, const LeafPair& lp, bool live) {
  const uint32_t lane = lane_id(), r = lane & 7u, shift = lane & ~7u;
  uint32_t sa = 0, sb = 0;
  if (live) {
    const uint2 bp = A.pairs[lp.pair];
    const BodyDev& BA = A.bodies[bp.x];
    const BodyDev& BB = A.bodies[bp.y];
    sa = BA.surf[lp.ta] & 15u;
    sb = BB.surf[lp.tb] & 15u;
    const BodyDev& Bd = r < 4u ? BA : BB;
    const uint32_t vi = Bd.tets[4 * (size_t)(r < 4u ? lp.ta : lp.tb) + (r & 3u)];
    V[r] = world_vertex(Bd, vi);
  }
  wave_lds_sync();
  if (live) {
    const f3* P = V + (r & 4u);
    const uint32_t f = r & 3u;
    U[r] = make_uplane(P[face_v(f, 0)], P[face_v(f, 1)], P[face_v(f, 2)]);
  }
  wave_lds_sync();
  bool sep = false;
  if (live)
    for (uint32_t ax = r; ax < 44u; ax += 8u) sep = sep || sat_axis(V, ax);
  const bool alive = live && (((__ballot(sep) >> shift) & 0xFFull) == 0ull);
  // face pairs c = 4a + b: lane r takes c = r and c = r + 8
  bool ok[2] = {false, false};
  float t[2] = {INFINITY, INFINITY};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t a = (r >> 2) + 2u * h, b = r & 3u;
    if (alive && ((sa >> a) & 1u) && ((sb >> b) & 1u)) {
      const UPlane plA = U[a], plB = U[4 + b];
      const f3 nA = unit3(plA.n), nB = unit3(plB.n);
      const float test = dot3(nA, nB);
      const float testA = dot3(plA.n, sub3(V[4 + b], plA.p));
      const float testB = dot3(plB.n, sub3(V[a], plB.p));
      ok[h] = test < 0.0f && !(testA < 0.0f) && !(testB < 0.0f);
      t[h] = test;
    }
  }
  float best = fminf(ok[0] ? t[0] : INFINITY, ok[1] ? t[1] : INFINITY);
#pragma unroll
  for (int d = 1; d < 8; d <<= 1) best = fminf(best, __shfl_xor(best, d, 64));
  const uint32_t m0 = (uint32_t)(__ballot(ok[0] && t[0] == best) >> shift) & 0xFFu;
  const uint32_t m1 = (uint32_t)(__ballot(ok[1] && t[1] == best) >> shift) & 0xFFu;
  wave_lds_sync();  // the group's LDS is rewritten by its next pair
  if (m0) return (__ffs((int)m0) - 1) >> 2;
  if (m1) return (8 + __ffs((int)m1) - 1) >> 2;
  return -1;
}

__device__ inline float wave_min_f(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ inline float wave_max_f(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ inline float readlane_f(float v, uint32_t l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// Narrow-phase output staged per block in LDS.  A reservation on one global
// counter per block iteration (tens of thousands per call for many surface
// pairs) serializes across the XCDs and sets the kernel's time; staged, a
// block reserves once per kStageC - kIterC contacts and once at its end.
constexpr uint32_t kStageC = 512, kStageR = 64;
constexpr uint32_t kIterC = 4u * (uint32_t)kMaxCand;  // most contacts one block iteration (4 waves) adds
struct NarrowStage {
  kdop_contact c[kStageC];
  NarrowRec r[kStageR];
  uint32_t nc, nr, base_c, base_r;
};

// Copy n staged records to out[base ..), as 32-bit words, dropping those at
// or past cap.
template <class T>
__device__ inline void stage_copy(const T* st, uint32_t n, T* __restrict__ out, uint32_t base, uint32_t cap) {
  static_assert(sizeof(T) % 4 == 0, "records of whole words");
  constexpr uint32_t W = sizeof(T) / 4;
  const uint32_t m = base < cap ? min(n, cap - base) : 0u;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
  uint32_t* dst = reinterpret_cast<uint32_t*>(out) + (size_t)base * W;
  for (uint32_t w = threadIdx.x; w < m * W; w += blockDim.x) dst[w] = src[w];
}

// Flush the staged contacts / records to global memory when forced (the
// block's end) or when the next iteration might not fit.  Every thread of
// the block calls it.
__device__ inline void narrow_stage_flush(const NarrowArgs& A, NarrowStage& S, bool force) {
  __syncthreads();
  const uint32_t nc = S.nc, nr = S.nr;
  const bool fc = nc != 0u && (force || nc > kStageC - kIterC);
  const bool fr = nr != 0u && (force || nr > kStageR - 4u);
  if (!fc && !fr) return;
  if (threadIdx.x == 0) {
    S.base_c = fc ? atomicAdd(A.c_fill, nc) : 0u;
    S.base_r = fr ? atomicAdd(A.r_fill, nr) : 0u;
    if (fc && S.base_c + nc > A.c_cap) atomicOr(A.overflow, 8u);
  }
  __syncthreads();
  if (fc) stage_copy(S.c, nc, A.contacts, S.base_c, A.c_cap);
  if (fr) stage_copy(S.r, nr, A.recs, S.base_r, A.r_cap);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (fc) S.nc = 0u;
    if (fr) S.nr = 0u;
  }
  __syncthreads();
}

// Contact points of one gated surface pair, one wave (contacts/geometry_
// contacts_tetrahedron_tetrahedron.h:543-681, as :1312-1345 calls it): lanes
// take the 56 contact candidates, each computed with the serial operation
// order; the serial loops' choices (first minimum, first maximum, a candidate
// dropped when within working precision of ANY earlier candidate) are made
// with ballots over lanes in that order, so the results are those of the
// serial code.  gate_word: P2 index | face << 30; live is wave-uniform.  The
// contacts and the pair's record go to the block's stage (room for them is
// left by narrow_stage_flush between iterations).
__device__ inline void narrow_contacts_wave(const NarrowArgs& A, NarrowLds& L, uint32_t gate_word, bool live, NarrowStage& S,
                                            unsigned long long& hits) {
  const uint32_t lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  LeafPair lp{0, 0, 0};
  const uint32_t face = gate_word >> 30;
  if (live) {
    const uint32_t g = gate_word;
    lp = A.p2[g & 0x3FFFFFFFu];
    const uint2 bp = A.pairs[lp.pair];
    if (lane < 8u) {
      const BodyDev& Bd = A.bodies[lane < 4u ? bp.x : bp.y];
      const uint32_t t = lane < 4u ? lp.ta : lp.tb;
      const uint32_t vi = Bd.tets[4 * (size_t)t + (lane & 3u)];
      L.v[lane] = world_vertex(Bd, vi);
    }
  }
  wave_lds_sync();
  // face planes (contacts ...tetrahedron.h:30-62)
  if (live && lane < 8u) {
    const f3* P = L.v + (lane & 4u);
    const uint32_t f = lane & 3u;
    L.u[lane] = make_uplane(P[face_v(f, 0)], P[face_v(f, 1)], P[face_v(f, 2)]);
  }
  wave_lds_sync();
  f3 nrm{0.f, 0.f, 0.f};
  bool ex = false;
  f3 q{0.f, 0.f, 0.f};
  uint8_t code = 0;
  if (live) {
    nrm = unit3(L.u[face].n);
    if (lane < (uint32_t)kMaxCand) ex = candidate(L, lane, &q, &code);
  }
  const unsigned long long exm = __ballot(ex);
  live = live && exm != 0ull;
  uint32_t nu = 0, ordinal = 0;
  bool keep = false;
  float depth = 0.0f;
  uint32_t cmin = 0, cmax = 0;
  if (live) {
  // penetration depth from the extreme projections onto the normal
    const float d = dot3(q, nrm);
    const bool has = ex && d == d;
    // serial: mn = FLT_MAX, mx = -FLT_MAX, first candidate attaining each
    // (strict comparisons); the value is read from that candidate's lane
    const float lo = wave_min_f(has ? d : FLT_MAX);
    const float hi = wave_max_f(has ? d : -FLT_MAX);
    const unsigned long long at_mn = __ballot(has && d == lo && d < FLT_MAX);
    const unsigned long long at_mx = __ballot(has && d == hi && -FLT_MAX < d);
    const uint32_t first = (uint32_t)(__ffsll(exm) - 1);
    const uint32_t lmin = at_mn ? (uint32_t)(__ffsll(at_mn) - 1) : first;
    const uint32_t lmax = at_mx ? (uint32_t)(__ffsll(at_mx) - 1) : first;
    const float mn = at_mn ? readlane_f(d, lmin) : FLT_MAX;
    const float mx = at_mx ? readlane_f(d, lmax) : -FLT_MAX;
    cmin = __builtin_amdgcn_readlane((uint32_t)code, lmin);
    cmax = __builtin_amdgcn_readlane((uint32_t)code, lmax);
    depth = mn - mx;
    // project onto the mid plane, drop points within working precision
    const f3 mid = mul3(mul3(nrm, mx + mn), 0.5f);
    const float s = dot3(nrm, sub3(q, mid));
    q = sub3(q, f3{s * nrm.x, s * nrm.y, s * nrm.z});
    const float wp = FLT_EPSILON * 10.0f;
    bool dup = false;
    unsigned long long rem = exm & (exm - 1ull) ? exm & ~(1ull << (63 - __clzll(exm))) : 0ull;  // all but the last
    while (rem) {
      const uint32_t src = (uint32_t)(__ffsll(rem) - 1);
      rem &= rem - 1ull;
      const f3 o{readlane_f(q.x, src), readlane_f(q.y, src), readlane_f(q.z, src)};
      if (lane > src) {
        const f3 dd = sub3(o, q);
        const float n1 = smax(smax(fabsf(dd.x), fabsf(dd.y)), fabsf(dd.z));
        if (n1 < wp) dup = true;
      }
    }
    keep = ex && !dup;
    const unsigned long long km = __ballot(keep);
    nu = (uint32_t)__popcll(km);
    ordinal = (uint32_t)__popcll(km & below);
  }
  uint32_t cb = 0, rb = 0;
  if (lane == 0u && nu) {
    cb = atomicAdd(&S.nc, nu);  // LDS atomics: this wave's places in the stage
    rb = atomicAdd(&S.nr, 1u);
  }
  cb = __shfl(cb, 0, 64);
  if (keep) {
    kdop_contact ct;
    ct.pair = lp.pair; ct.tet_a = lp.ta; ct.tet_b = lp.tb; ct.ordinal = ordinal;
    ct.p[0] = q.x; ct.p[1] = q.y; ct.p[2] = q.z;
    ct.n[0] = nrm.x; ct.n[1] = nrm.y; ct.n[2] = nrm.z;
    ct.depth = depth;
    S.c[cb + ordinal] = ct;
  }
  if (lane == 0u && nu) {
    ++hits;
    NarrowRec rec;
    rec.pair = lp.pair; rec.ta = lp.ta; rec.tb = lp.tb;
    rec.face = (uint8_t)face; rec.nuniq = (uint8_t)nu; rec.cmin = (uint8_t)cmin; rec.cmax = (uint8_t)cmax;
    rec.depth = depth; rec.n[0] = nrm.x; rec.n[1] = nrm.y; rec.n[2] = nrm.z;
    S.r[rb] = rec;
  }
  wave_lds_sync();  // this wave's LDS is rewritten by its next pair
}

// The launch's control words -- the fills of the round's segments (A.pending:
// the narrow phase runs only when the persistent rounds left nothing), the
// done flag and the surface-pair count -- and this group's first pair entry
// (slot `first`, read before the count is known; used only below it), all in
// one round trip.  Returns false when the launch has nothing to do; *n: the
// pairs to take.
__device__ inline bool narrow_start(const NarrowArgs& A, uint32_t first, uint32_t* n, LeafPair* lp) {
  uint32_t pv = 0u, dn = 1u;
  if (A.pending) pv = A.pending[(size_t)lane_id() * A.pending_stride];
  if (A.done) dn = *A.done;
  const uint32_t fill = *A.p2_fill;
  *lp = LeafPair{0u, 0u, 0u};
  if (first < A.p2_cap) *lp = A.p2[first];
  if (__any(pv != 0u) || dn == 0u) return false;
  *n = min(fill, A.p2_cap);
  return true;
}

// Narrow phase for few surface pairs: one wave per pair.  Lanes 0..43 test the
// SAT axes and lanes 0..15 the face pairs (the serial loops' results by
// ballots), then narrow_contacts_wave on the same wave.
__global__ __launch_bounds__(256) void k_narrow_wave(NarrowArgs A) {
  __shared__ NarrowLds lds[4];
  __shared__ NarrowStage S;
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t n = 0;
  LeafPair lp_next;
  if (!narrow_start(A, blockIdx.x * 4u + w, &n, &lp_next)) return;
  NarrowLds& L = lds[w];
  unsigned long long hits = 0;
  if (threadIdx.x == 0) S.nc = S.nr = 0u;
  __syncthreads();
  for (uint32_t base = blockIdx.x * 4u; base < n; base += gridDim.x * 4u) {
    const uint32_t i = base + w;
    bool live = i < n;  // wave-uniform
    const LeafPair lp = lp_next;  // loaded an iteration ahead
    if (i + gridDim.x * 4u < n) lp_next = A.p2[i + gridDim.x * 4u];
    uint32_t sa = 0, sb = 0;
    if (live) {
      const uint2 bp = A.pairs[lp.pair];
      const BodyDev& BA = A.bodies[bp.x];
      const BodyDev& BB = A.bodies[bp.y];
      sa = BA.surf[lp.ta] & 15u;
      sb = BB.surf[lp.tb] & 15u;
      if (lane < 8u) {
        const BodyDev& Bd = lane < 4u ? BA : BB;
        const uint32_t vi = Bd.tets[4 * (size_t)(lane < 4u ? lp.ta : lp.tb) + (lane & 3u)];
        L.v[lane] = world_vertex(Bd, vi);
      }
    }
    wave_lds_sync();
    if (live && lane < 8u) {
      const f3* P = L.v + (lane & 4u);
      const uint32_t f = lane & 3u;
      L.u[lane] = make_uplane(P[face_v(f, 0)], P[face_v(f, 1)], P[face_v(f, 2)]);
    }
    wave_lds_sync();
    const bool sep = live && lane < 44u && sat_axis(L.v, lane);
    live = live && !__any(sep);
    bool ok = false;
    float test = 0.0f;
    if (live && lane < 16u) {
      const uint32_t a = lane >> 2, b = lane & 3u;
      if (((sa >> a) & 1u) && ((sb >> b) & 1u)) {
        const UPlane plA = L.u[a], plB = L.u[4 + b];
        const f3 nA = unit3(plA.n), nB = unit3(plB.n);
        test = dot3(nA, nB);
        const float testA = dot3(plA.n, sub3(L.v[4 + b], plA.p));
        const float testB = dot3(plB.n, sub3(L.v[a], plB.p));
        ok = test < 0.0f && !(testA < 0.0f) && !(testB < 0.0f);
      }
    }
    const float best = wave_min_f(ok ? test : INFINITY);
    const unsigned long long pick = __ballot(ok && test == best);
    const uint32_t face = pick ? (uint32_t)(__ffsll(pick) - 1) >> 2 : 0u;
    wave_lds_sync();
    narrow_contacts_wave(A, L, i | (face << 30), live && pick != 0ull, S, hits);
    narrow_stage_flush(A, S, false);
  }
  narrow_stage_flush(A, S, true);
  wave_add64(&stripe_of(A.stripes, wave_global())->narrow_pairs, hits);
}

// Narrow phase for many surface pairs, pass 1: the SAT gate and face pick
// with 8 lanes per pair (gate_group, 32 pairs per block at a time); the
// survivors go to the gate list (P2 index | face << 30).
// Survivors are staged in LDS and written with one reservation per
// kGateStage - 32 of them (and one at the block's end).
constexpr uint32_t kGateStage = 1024;
__global__ __launch_bounds__(256) void k_narrow_gate(NarrowArgs A) {
  __shared__ f3 sv[32][8];
  __shared__ UPlane su[32][8];
  __shared__ uint32_t gst[kGateStage];
  __shared__ uint32_t gn, gbase;
  const uint32_t lane = lane_id(), grp = threadIdx.x >> 3;
  uint32_t n = 0;
  LeafPair lp_next;
  if (!narrow_start(A, blockIdx.x * 32u + grp, &n, &lp_next)) return;
  if (threadIdx.x == 0) gn = 0u;
  __syncthreads();
  for (uint32_t base = blockIdx.x * 32u;; base += gridDim.x * 32u) {
    const bool more = base < n;  // block-uniform
    if (more) {
      const uint32_t i = base + grp;
      const LeafPair lp = lp_next;  // loaded an iteration ahead
      if (i + gridDim.x * 32u < n) lp_next = A.p2[i + gridDim.x * 32u];
      const int face = gate_group(A, sv[grp], su[grp], lp, i < n);
      if ((lane & 7u) == 0u && face >= 0) gst[atomicAdd(&gn, 1u)] = i | ((uint32_t)face << 30);  // <= 32 per iteration
    }
    __syncthreads();
    const uint32_t m = gn;
    if (m != 0u && (!more || m > kGateStage - 32u)) {
      if (threadIdx.x == 0) gbase = atomicAdd(A.g_fill, m);
      __syncthreads();
      for (uint32_t k = threadIdx.x; k < m; k += blockDim.x)
        if (gbase + k < A.p2_cap) A.gate[gbase + k] = gst[k];  // the gate list holds at most n <= P2 capacity words
      __syncthreads();
      if (threadIdx.x == 0) gn = 0u;
      __syncthreads();
    }
    if (!more) break;
  }
}

// Narrow phase for many surface pairs, pass 2: contact points of the gated
// pairs, one wave each.
__global__ __launch_bounds__(256) void k_narrow_gated(NarrowArgs A) {
  __shared__ NarrowLds lds[4];
  __shared__ NarrowStage S;
  const uint32_t n = *A.g_fill;
  const uint32_t w = threadIdx.x >> 6;
  unsigned long long hits = 0;
  if (threadIdx.x == 0) S.nc = S.nr = 0u;
  __syncthreads();
  for (uint32_t base = blockIdx.x * 4u; base < n; base += gridDim.x * 4u) {
    const uint32_t k = base + w;
    narrow_contacts_wave(A, lds[w], k < n ? A.gate[k] : 0u, k < n, S, hits);
    narrow_stage_flush(A, S, false);
  }
  narrow_stage_flush(A, S, true);
  wave_add64(&stripe_of(A.stripes, wave_global())->narrow_pairs, hits);
}

}  // namespace kdop_amd
