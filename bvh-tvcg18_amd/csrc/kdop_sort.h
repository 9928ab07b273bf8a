// kdop_sort.h -- deterministic device sort and its two users: the contact
// report order and the Morton reorder of a tetrahedral mesh.
//
// Sort: every key type has a strict total order (its source index breaks
// every tie), so the result does not depend on how the work is split.  Tiles
// of kSortTile keys are sorted in LDS by a bitonic network (k_sort_tiles),
// then runs of width w are merged pairwise (k_merge_pass): a key's place in
// the merged run is its place in its own run plus the number of keys of the
// other run below it (a binary search), so each pass is one gather-free
// scatter.  Exclusive scans of uint32 flags (k_scan_*) serve the reorder.
#pragma once

#include <stdint.h>

#include "kdop_amd.h"
#include "kdop_device.h"
#include "kdop_math.h"

namespace kdop_amd {

constexpr uint32_t kSortTile = 1024, kSortThreads = 256;

// Morton code of a tetrahedron (kdop_mesh_reorder.h:262-285), ties by index.
struct MortonKey {
  unsigned long long code;
  uint32_t idx, pad;
  __device__ bool operator<(const MortonKey& o) const { return code != o.code ? code < o.code : idx < o.idx; }
  __device__ static MortonKey top() { return MortonKey{~0ull, ~0u, 0u}; }
};

// The sequential reference's report order (kdop_tandem_traversal.h:160-196,
// 18-156): per body pair, the branch chunks of A and B in row-major order,
// then the depth-first descent inside the chunk pair (children (A0,B0),
// (A0,B1), (A1,B0), (A1,B1) when both nodes are internal, the internal side's
// two children when one is a leaf), then the order the contact generator
// produced the points in (the contact's ordinal).
struct ContactSeqKey {
  uint32_t pair, ca, cb;
  unsigned long long path;
  uint32_t ord, idx;
  __device__ bool operator<(const ContactSeqKey& o) const {
    if (pair != o.pair) return pair < o.pair;
    if (ca != o.ca) return ca < o.ca;
    if (cb != o.cb) return cb < o.cb;
    if (path != o.path) return path < o.path;
    if (ord != o.ord) return ord < o.ord;
    return idx < o.idx;
  }
  __device__ static ContactSeqKey top() { return ContactSeqKey{~0u, ~0u, ~0u, ~0ull, ~0u, ~0u}; }
};

// Descent path to the leaf pair (ja, jb) of two branch chunks of na and nb
// tets (median split of the local index range, as the chunk trees are built):
// one base-4 digit per step, 2 ca + cb while both nodes are internal, else the
// internal side's choice, left-aligned in 2 * kMaxPathSteps bits so that
// integer order is the descent's order (two leaf pairs of one chunk pair
// share their steps up to the node pair where they part, and no path is a
// prefix of another).  A path has as many steps as the deeper chunk tree,
// at most ceil(log2 L) <= 30 for the node counts the library takes.
constexpr uint32_t kMaxPathSteps = 32;
__device__ inline unsigned long long descent_path(uint32_t ja, uint32_t na, uint32_t jb, uint32_t nb) {
  uint32_t loA = 0, hiA = na - 1u, loB = 0, hiB = nb - 1u, steps = 0;
  unsigned long long key = 0;
  while (loA < hiA || loB < hiB) {
    const bool ia = loA < hiA, ib = loB < hiB;
    uint32_t ca = 0, cb = 0;
    if (ia) {
      const uint32_t mid = (loA + hiA) >> 1;
      ca = ja > mid ? 1u : 0u;
      if (ca) loA = mid + 1u; else hiA = mid;
    }
    if (ib) {
      const uint32_t mid = (loB + hiB) >> 1;
      cb = jb > mid ? 1u : 0u;
      if (cb) loB = mid + 1u; else hiB = mid;
    }
    key = (key << 2) | ((ia && ib) ? 2u * ca + cb : (ia ? ca : cb));
    ++steps;
  }
  return steps ? key << (2u * (kMaxPathSteps - steps)) : 0ull;
}

// A float as a uint32 whose unsigned order is the float order, with -0 and +0
// mapped to one value (they compare equal as floats) and NaNs placed past the
// infinities, so a key built from it always has a strict total order (a float
// comparator with a NaN in it has none, and the rank merge would then leave
// slots unwritten).  For every non-NaN pair, a < b as floats iff ordered(a) <
// ordered(b), and a == b iff ordered(a) == ordered(b).
__device__ inline uint32_t ordered_bits(float f) {
  const uint32_t u = f == 0.0f ? 0u : __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ordered_float(uint32_t o) {  // the float back (+0 for either zero)
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// The OpenCL path's report order: qsort by compare_kernel_contact_points
// (cl/kdop_cl_tandem_traversal.h:61-104, applied at :1772-1788): test pair,
// then p, n (component-wise) and depth with float comparisons, equal when
// neither is below the other; the source index breaks the remaining ties
// (qsort leaves them in no particular order).
struct ContactClKey {
  uint32_t pair;
  uint32_t v[7];  // ordered_bits of p[3], n[3], depth
  uint32_t idx;
  __device__ bool operator<(const ContactClKey& o) const {
    if (pair != o.pair) return pair < o.pair;
#pragma unroll
    for (int i = 0; i < 7; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return idx < o.idx;
  }
  __device__ static ContactClKey top() {
    ContactClKey k;
    k.pair = ~0u;
#pragma unroll
    for (int i = 0; i < 7; ++i) k.v[i] = ~0u;
    k.idx = ~0u;
    return k;
  }
};

// PROX contact reduction (SIMULATION/PROX/PROX/include/prox_collision_detection.h:
// 340-345: std::sort, then std::unique) over prox::ContactPoint
// (prox_contact_point.h).  Its operator< (:140-190) orders by body i, body j,
// then p.y, p.x, p.z, n.y, n.x, n.z, depth with float comparisons; the source
// position breaks the remaining ties, which leaves points equal in every
// compared field (std::sort leaves them in no particular order; they can
// differ only in the sign of a zero).  w: body i, body j, then ordered_bits
// of the seven floats in that order.
struct ProxKey {
  uint32_t w[9];
  uint32_t idx;
  __device__ bool operator<(const ProxKey& o) const {
#pragma unroll
    for (int i = 0; i < 9; ++i)
      if (w[i] != o.w[i]) return w[i] < o.w[i];
    return idx < o.idx;
  }
  __device__ static ProxKey top() {
    ProxKey k;
#pragma unroll
    for (int i = 0; i < 9; ++i) k.w[i] = ~0u;
    k.idx = ~0u;
    return k;
  }
};

// Sort the tile of kSortTile keys at blockIdx.x * kSortTile in place (LDS
// bitonic network; the tail of the last tile is padded with top()).
template <class Key>
__global__ __launch_bounds__(kSortThreads) void k_sort_tiles(Key* __restrict__ keys, uint32_t n) {
  __shared__ Key s[kSortTile];
  const uint32_t base = blockIdx.x * kSortTile, tid = threadIdx.x;
  for (uint32_t i = tid; i < kSortTile; i += kSortThreads) s[i] = base + i < n ? keys[base + i] : Key::top();
  __syncthreads();
  for (uint32_t k = 2; k <= kSortTile; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = tid; t < kSortTile / 2; t += kSortThreads) {
        const uint32_t i = 2u * t - (t & (j - 1u)), p = i + j;  // i has bit j clear
        const bool up = (i & k) == 0u;
        const Key a = s[i], b = s[p];
        if ((b < a) == up) { s[i] = b; s[p] = a; }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = tid; i < kSortTile; i += kSortThreads)
    if (base + i < n) keys[base + i] = s[i];
}

// Merge the sorted runs [2r w, (2r+1) w) and [(2r+1) w, (2r+2) w) of `in`
// into `out` (one thread per key; a run without a partner is copied).
template <class Key>
__global__ __launch_bounds__(256) void k_merge_pass(const Key* __restrict__ in, Key* __restrict__ out, uint32_t n, uint32_t w) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t r = g / w, p0 = (r ^ 1u) * w;
  const Key k = in[g];
  if (p0 >= n) { out[g] = k; return; }
  uint32_t lo = p0, hi = min(p0 + w, n);
  while (lo < hi) {  // partner keys below k
    const uint32_t mid = (lo + hi) >> 1;
    if (in[mid] < k) lo = mid + 1u; else hi = mid;
  }
  out[(r & ~1u) * w + (g - r * w) + (lo - p0)] = k;
}

// ---------------------------------------------------------------- scans
// Exclusive scan of n uint32 in three launches: per-block scans of
// kScanBlock values with the block totals, a scan of the totals by one
// block, then the block offsets added.  total: the sum of all values.
constexpr uint32_t kScanBlock = 1024;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  // lds: kScanBlock / 64 + 1 words
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63u) lds[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t i = 0; i < blockDim.x / 64u; ++i) {
      const uint32_t t = lds[i];
      lds[i] = acc;
      acc += t;
    }
    lds[blockDim.x / 64u] = acc;
  }
  __syncthreads();
  const uint32_t r = lds[w] + x - v;
  *total = lds[blockDim.x / 64u];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                                                          uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t lds[kScanBlock / 64 + 1];
  const uint32_t g = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t v = g < n ? in[g] : 0u;
  uint32_t tot;
  const uint32_t r = block_exclusive_scan(v, lds, &tot);
  if (g < n) out[g] = r;
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ total) {
  __shared__ uint32_t lds[kScanBlock / 64 + 1];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += kScanBlock) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t tot;
    const uint32_t r = block_exclusive_scan(v, lds, &tot);
    if (i < nb) sums[i] = carry + r;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums) {
  const uint32_t g = blockIdx.x * kScanBlock + threadIdx.x;
  if (g < n) out[g] += sums[blockIdx.x];
}

// ------------------------------------------------------------- contacts
// pairs: body ids of the traversal's pairs; a tet's branch chunk is tet / L
// (chunk c holds tets [c L, c L + L) of the body, the last one fewer).
__device__ inline uint32_t chunk_size(const BodyDev& B, uint32_t c) { return c + 1u < B.nbranch ? B.L : B.nt - c * B.L; }
__global__ void k_contact_keys_seq(const kdop_contact* __restrict__ c, uint32_t n, const uint2* __restrict__ pairs,
                                   const BodyDev* __restrict__ bodies, ContactSeqKey* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const kdop_contact& x = c[i];
  const uint2 bp = pairs[x.pair];
  const BodyDev& A = bodies[bp.x];
  const BodyDev& B = bodies[bp.y];
  const uint32_t ca = x.tet_a / A.L, cb = x.tet_b / B.L;
  const unsigned long long path = descent_path(x.tet_a - ca * A.L, chunk_size(A, ca), x.tet_b - cb * B.L, chunk_size(B, cb));
  keys[i] = ContactSeqKey{x.pair, ca, cb, path, x.ordinal, i};
}

__global__ void k_contact_keys_cl(const kdop_contact* __restrict__ c, uint32_t n, ContactClKey* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const kdop_contact& x = c[i];
  ContactClKey k;
  k.pair = x.pair;
  k.v[0] = ordered_bits(x.p[0]); k.v[1] = ordered_bits(x.p[1]); k.v[2] = ordered_bits(x.p[2]);
  k.v[3] = ordered_bits(x.n[0]); k.v[4] = ordered_bits(x.n[1]); k.v[5] = ordered_bits(x.n[2]);
  k.v[6] = ordered_bits(x.depth);
  k.idx = i;
  keys[i] = k;
}

template <class Key>
__global__ void k_contact_permute(const kdop_contact* __restrict__ in, const Key* __restrict__ keys, uint32_t n,
                                  kdop_contact* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[keys[i].idx];
}

// ----------------------------------------------------- contact reduction
__device__ inline ProxKey prox_key(uint32_t bi, uint32_t bj, const float* p, const float* n, float depth, uint32_t idx) {
  ProxKey k;
  k.w[0] = bi; k.w[1] = bj;
  k.w[2] = ordered_bits(p[1]); k.w[3] = ordered_bits(p[0]); k.w[4] = ordered_bits(p[2]);
  k.w[5] = ordered_bits(n[1]); k.w[6] = ordered_bits(n[0]); k.w[7] = ordered_bits(n[2]);
  k.w[8] = ordered_bits(depth);
  k.idx = idx;
  return k;
}
// Contacts of the last traversal: body i and body j are the pair's bodies A
// and B (ContactCallbackFunctor, prox_collision_detection.h:38-110).
__global__ void k_prox_keys_pairs(const kdop_contact* __restrict__ c, uint32_t n, const uint2* __restrict__ pairs, ProxKey* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const kdop_contact& x = c[i];
  const uint2 bp = pairs[x.pair];
  keys[i] = prox_key(bp.x, bp.y, x.p, x.n, x.depth, i);
}
__global__ void k_prox_keys_points(const kdop_prox_contact* __restrict__ c, uint32_t n, ProxKey* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const kdop_prox_contact& x = c[i];
  keys[i] = prox_key(x.body_i, x.body_j, x.p, x.n, x.depth, i);
}

// tiny::working_precision<float>() (tiny_precision.h:21-24): epsilon * 10.
constexpr float kProxPrecision = FLT_EPSILON * 10.0f;

// ContactPoint::operator== (prox_contact_point.h:91-138): the same bodies and
// every compared float within the working precision (|difference| > it fails).
__device__ inline bool prox_equal(const ProxKey& a, const ProxKey& b) {
  if (a.w[0] != b.w[0] || a.w[1] != b.w[1]) return false;
#pragma unroll
  for (int i = 2; i < 9; ++i)
    if (fabsf(ordered_float(b.w[i]) - ordered_float(a.w[i])) > kProxPrecision) return false;
  return true;
}
// b follows a in sorted order and can equal no point before it: other bodies
// (the body pair's points are contiguous), or p.y more than the precision
// above a's (the points of one body pair are sorted by p.y, and a correctly
// rounded difference only grows as the subtrahend falls, so every earlier
// point of the pair is at least as far below).
__device__ inline bool prox_break(const ProxKey& a, const ProxKey& b) {
  return a.w[0] != b.w[0] || a.w[1] != b.w[1] || fabsf(ordered_float(b.w[2]) - ordered_float(a.w[2])) > kProxPrecision;
}

// std::unique as libstdc++ runs it: a point is kept unless it equals the last
// point kept (not the point before it: the tolerance makes equality
// intransitive, so the difference shows).  A point after a break is kept
// whatever came before, and what follows it depends on nothing before it, so
// the sorted points split at the breaks into segments handled independently:
// the thread at a segment's start walks its segment and writes keep (1/0) for
// each of its points.  Segments are short unless many points share p.y within
// the precision.
__global__ __launch_bounds__(256) void k_prox_unique(const ProxKey* __restrict__ keys, uint32_t n, uint32_t* __restrict__ keep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ProxKey prev = keys[i];
  if (i > 0u && !prox_break(keys[i - 1u], prev)) return;
  ProxKey last = prev;
  keep[i] = 1u;
  for (uint32_t j = i + 1u; j < n; ++j) {
    const ProxKey k = keys[j];
    if (prox_break(prev, k)) break;
    const bool same = prox_equal(last, k);
    keep[j] = same ? 0u : 1u;
    if (!same) last = k;
    prev = k;
  }
}

// out[off[i]] = in[keys[i].idx] for the kept points, in sorted order.
template <class Rec>
__global__ void k_prox_compact(const Rec* __restrict__ in, const ProxKey* __restrict__ keys, const uint32_t* __restrict__ keep,
                               const uint32_t* __restrict__ off, uint32_t n, Rec* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i]) out[off[i]] = in[keys[i].idx];
}

// -------------------------------------------------------- Morton reorder
// kdop_mesh_reorder.h:172-353.  Centroid per tet: the float sum of the four
// coordinates times 0.25 in double, stored as float (:193-221); the minimum
// starts at FLT_MAX and the maximum at FLT_MIN, the smallest positive
// normal float, as numeric_limits<float>::min() gives (:197-199).
struct ReorderParams {
  float lo[3], hi[3];
  int scale;
  uint32_t pad;
};

__device__ inline float centroid_axis(const float* __restrict__ A, const uint32_t* v) {
  return (float)(0.25 * (double)(A[v[0]] + A[v[1]] + A[v[2]] + A[v[3]]));
}

// Per block: min and max of the block's centroids on each axis (part: 6 floats per block).
__global__ __launch_bounds__(256) void k_reorder_bounds(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
                                                      const uint32_t* __restrict__ tets, uint32_t nt, float* __restrict__ part) {
  __shared__ float red[6][4];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float c[6] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN, FLT_MIN};
  if (t < nt) {
    const uint4 q = *reinterpret_cast<const uint4*>(tets + 4 * (size_t)t);
    const uint32_t v[4] = {q.x, q.y, q.z, q.w};
    const float x = centroid_axis(X, v), y = centroid_axis(Y, v), z = centroid_axis(Z, v);
    c[0] = x; c[1] = y; c[2] = z; c[3] = smax(FLT_MIN, x); c[4] = smax(FLT_MIN, y); c[5] = smax(FLT_MIN, z);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const float o = __shfl_xor(c[a], d, 64);
      c[a] = a < 3 ? smin(c[a], o) : smax(c[a], o);
    }
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 6; ++a) red[a][w] = c[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = (int)threadIdx.x;
    float r = red[a][0];
    for (uint32_t i = 1; i < blockDim.x / 64u; ++i) r = a < 3 ? smin(r, red[a][i]) : smax(r, red[a][i]);
    part[6 * (size_t)blockIdx.x + a] = r;
  }
}

// One block: fold the per-block bounds and derive the common integer scale
// (kdop_mesh_reorder.h:226-240): scale = int((2^21 - 1) / (max - min)) per
// axis, the smallest of the three; an axis whose centroids are all equal
// gives INT_MAX instead of the reference's undefined conversion.
__device__ inline int reorder_axis_scale(float lo, float hi) {
  const float s = 2097151.0f / (hi - lo);
  return (s < 2147483647.0f) ? (int)s : 2147483647;
}
__global__ __launch_bounds__(256) void k_reorder_scale(const float* __restrict__ part, uint32_t nb, ReorderParams* __restrict__ P) {
  __shared__ float red[6][4];
  float c[6] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN, FLT_MIN};
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int a = 0; a < 6; ++a) c[a] = a < 3 ? smin(c[a], part[6 * (size_t)b + a]) : smax(c[a], part[6 * (size_t)b + a]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const float o = __shfl_xor(c[a], d, 64);
      c[a] = a < 3 ? smin(c[a], o) : smax(c[a], o);
    }
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 6; ++a) red[a][w] = c[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    float r[6];
    for (int a = 0; a < 6; ++a) {
      r[a] = red[a][0];
      for (uint32_t i = 1; i < blockDim.x / 64u; ++i) r[a] = a < 3 ? smin(r[a], red[a][i]) : smax(r[a], red[a][i]);
    }
    int s = reorder_axis_scale(r[0], r[3]);
    s = min(s, reorder_axis_scale(r[1], r[4]));
    s = min(s, reorder_axis_scale(r[2], r[5]));
    ReorderParams p;
    p.lo[0] = r[0]; p.lo[1] = r[1]; p.lo[2] = r[2];
    p.hi[0] = r[3]; p.hi[1] = r[4]; p.hi[2] = r[5];
    p.scale = s;
    p.pad = 0;
    *P = p;
  }
}

// Morton key per tet: each centroid coordinate to int((c - min) * scale),
// kept as a float like the reference's attribute, then its 21 low bits
// interleaved x, y, z from bit 0 (kdop_mesh_reorder.h:244-285).
__device__ inline unsigned long long spread3(unsigned long long x) {
  x &= 0x1FFFFFull;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ inline unsigned long long quantise(float c, float lo, int scale) {
  return (unsigned long long)(float)(int)((c - lo) * (float)scale);
}
__global__ __launch_bounds__(256) void k_reorder_codes(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
                                                     const uint32_t* __restrict__ tets, uint32_t nt, const ReorderParams* __restrict__ P,
                                                     MortonKey* __restrict__ keys) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const ReorderParams p = *P;
  const uint4 q = *reinterpret_cast<const uint4*>(tets + 4 * (size_t)t);
  const uint32_t v[4] = {q.x, q.y, q.z, q.w};
  const unsigned long long x = quantise(centroid_axis(X, v), p.lo[0], p.scale);
  const unsigned long long y = quantise(centroid_axis(Y, v), p.lo[1], p.scale);
  const unsigned long long z = quantise(centroid_axis(Z, v), p.lo[2], p.scale);
  keys[t] = MortonKey{spread3(x) | (spread3(y) << 1) | (spread3(z) << 2), t, 0u};
}

// First touch of each vertex in the sorted tet order: the smallest slot
// 4 i + q (i: sorted position, q: vertex of the tet) that names it.
__global__ __launch_bounds__(256) void k_reorder_first(const MortonKey* __restrict__ keys, const uint32_t* __restrict__ tets, uint32_t nt,
                                                     uint32_t* __restrict__ first) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 4u * nt) return;
  const uint32_t v = tets[4 * (size_t)keys[s >> 2].idx + (s & 3u)];
  atomicMin(&first[v], s);
}

// flag[s] = slot s is its vertex's first touch; unused[v] = vertex v is named by no tet.
__global__ __launch_bounds__(256) void k_reorder_flags(const MortonKey* __restrict__ keys, const uint32_t* __restrict__ tets, uint32_t nt,
                                                     const uint32_t* __restrict__ first, uint32_t nv, uint32_t* __restrict__ flag,
                                                     uint32_t* __restrict__ unused) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < 4u * nt) {
    const uint32_t v = tets[4 * (size_t)keys[s >> 2].idx + (s & 3u)];
    flag[s] = first[v] == s ? 1u : 0u;
  }
  if (s < nv) unused[s] = first[s] == 0xFFFFFFFFu ? 1u : 0u;
}

// New vertex number: the rank of its first touch among first touches, and
// vertices no tet names after them in input order (meshgen's rule for them).
__device__ inline uint32_t new_vertex(uint32_t v, const uint32_t* first, const uint32_t* flag_scan, const uint32_t* unused_scan,
                                      uint32_t nused) {
  const uint32_t f = first[v];
  return f != 0xFFFFFFFFu ? flag_scan[f] : nused + unused_scan[v];
}
__global__ __launch_bounds__(256) void k_reorder_tets(const MortonKey* __restrict__ keys, const uint32_t* __restrict__ tets, uint32_t nt,
                                                    const uint32_t* __restrict__ first, const uint32_t* __restrict__ flag_scan,
                                                    uint32_t* __restrict__ tets_o) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 4u * nt) return;
  const uint32_t v = tets[4 * (size_t)keys[s >> 2].idx + (s & 3u)];
  tets_o[s] = flag_scan[first[v]];
}
__global__ __launch_bounds__(256) void k_reorder_verts(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
                                                     uint32_t nv, const uint32_t* __restrict__ first, const uint32_t* __restrict__ flag_scan,
                                                     const uint32_t* __restrict__ unused_scan, const uint32_t* __restrict__ nused,
                                                     float* __restrict__ Xo, float* __restrict__ Yo, float* __restrict__ Zo) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const uint32_t o = new_vertex(v, first, flag_scan, unused_scan, *nused);
  Xo[o] = X[v];
  Yo[o] = Y[v];
  Zo[o] = Z[v];
}

}  // namespace kdop_amd
