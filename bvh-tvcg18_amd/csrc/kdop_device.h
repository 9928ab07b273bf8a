// kdop_device.h -- device-side data layout shared by the kDOP kernels.
//
// Node record (AoS, one per tree node slot, rec_stride(K) floats: the power
// of two >= K + 2, i.e. 32 B for K = 6, 64 B for K = 8..14, 128 B for K =
// 18..26, so a record never straddles a 128-byte cache line):
//   float vol[K]   slab k: vol[2k] = lower, vol[2k+1] = upper
//   uint32 m0      internal: left child (global index)  | flags
//                  tet leaf: tet index (body local)      | flags
//   uint32 m1      internal: right child (global index) | F_ROOT when branch root
//                  tet leaf: body index | F_SURF when the tet has a surface face
//                            | F_ROOT when the leaf is a one-tet chunk's root
// Flags of m0: F_LEAF (tet leaf), F_BRANCH (node lies in a branch chunk).
// Super-chunk leaves of the reference tree (copies of a child chunk's root
// volume) are not stored: a super node points straight at the child chunk's
// root, resolved through one-element chunks.
// Slots: a chunk of stride 2L (branch) or L (super) keeps slot 0 free and its
// BFS node j in slot j + 1, so siblings (BFS 2j+1, 2j+2, or any later pair the
// median split appends) start on an even slot and share a line where records
// are <= 64 bytes.  Free slots hold meta words 0xFFFFFFFF.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kdop_amd {

constexpr uint32_t F_LEAF = 0x80000000u;
constexpr uint32_t F_BRANCH = 0x40000000u;
constexpr uint32_t F_ROOT = 0x80000000u;
constexpr uint32_t F_SURF = 0x40000000u;
constexpr uint32_t IDX_MASK = 0x3FFFFFFFu;
constexpr int kMaxLevels = 24;

__host__ __device__ constexpr int rec_stride(int K) { return K + 2 <= 8 ? 8 : (K + 2 <= 16 ? 16 : 32); }

// Per-body descriptor, device resident (one entry per body of the context).
struct BodyDev {
  uint32_t node_base;  // global index of the root record (body-local node 0)
  uint32_t nnodes;
  uint32_t nt, nv;
  uint32_t top_base;   // offset into the context's table of top-level chunk roots
  uint32_t ntop;
  uint32_t nbranch;    // branch chunks C_A
  uint32_t H, L;
  uint32_t branch_base;  // global index of the first branch-chunk node
  uint32_t depth;        // longest root-record-to-leaf path
  uint32_t super_base;   // H >= 2: global index of the first node slot of level H-2
  uint32_t super3_base;  // H >= 3: global index of the first node slot of level H-3
  const uint32_t* tets;
  const uint8_t* surf;
  const float* X0;
  const float* Y0;
  const float* Z0;
  const float4* W0;  // the same rest positions as (x, y, z, 0): one 16-byte load per vertex for the refit
  float* X;
  float* Y;
  float* Z;
  float4* W;  // the same world vertices as (x, y, z, 0): one 16-byte load per vertex for the narrow phase and the ray/sphere tests
};

// Topology parameters of one body (kernel argument).
struct TopoParams {
  uint32_t node_base, H, L, nt, body;
  uint32_t C[kMaxLevels];          // chunks per level (0 = top, H-1 = branches)
  uint32_t E[kMaxLevels];          // elements per level (tets at H-1, chunks of level h+1 above)
  uint32_t level_base[kMaxLevels]; // body-local node index of the level's first chunk
};

// Direction table as a kernel argument (host computed, bit-exact).
struct DirTable {
  float d[39];
};

// Work item of the traversal: a node pair known to overlap, with what its
// expansion needs already resolved by the producer (which held both records):
// per side either I_EXP | first child, second child (expand this side) or the
// node itself (a0 = node, a1 unused).  pair: index into the call's pair list.
constexpr uint32_t I_EXP = 0x80000000u;
struct Item {
  uint32_t a0, a1, b0, b1, pair;
};

// Leaf pair output: body-pair index and the two tetrahedra.
struct LeafPair {
  uint32_t pair, ta, tb;
};

// Narrow-phase record per surface pair that produced contacts (adjoint input).
struct NarrowRec {
  uint32_t pair, ta, tb;
  uint8_t face, nuniq, cmin, cmax;
  float depth;
  float n[3];
};

// Device counters of one call, striped: a kernel adds its per-wave sums to
// stripe (global wave index mod kStripes).  Agent-scope atomics on one word
// serialize across the eight XCDs (they are performed beyond the per-XCD
// L2), so thousands of waves adding to one counter cost more than the
// traversal itself; 64 stripes on separate 256-byte lines remove that hot
// spot.  k_sum_counters folds the stripes into `total`.
constexpr int kStripes = 64;
struct alignas(256) CounterStripe {
  unsigned long long tests;          // node-pair tests executed
  unsigned long long descent;        // tests below branch-root pairs
  unsigned long long branch_overlaps;
  unsigned long long p1;             // overlapping tet pairs
  unsigned long long narrow_pairs;   // surface pairs with contacts
  unsigned long long branch_pairs;   // C_A*C_B over body pairs whose roots overlap
  unsigned long long root_tests;
  unsigned long long iters;          // persistent: wave loop iterations
  unsigned long long lanes;          // persistent: items expanded
  unsigned long long grabs;          // persistent: frontier chunks taken
  unsigned long long ptests;         // persistent: node-pair tests made by k_persist
  unsigned long long handoffs;       // persistent: items the tail hand-off passed to the next round
  unsigned long long records;        // persistent: node records loaded by the expansions (one per record and item)
  unsigned long long max_iters;      // persistent: most iterations of one wave (max, not sum; last field)
};
constexpr int kStripeFields = 14;

// Segments of the persistent traversal's frontier: each round's input is
// split into kSegs segments with their own chunk counters (one lane of a
// wave per segment).
constexpr int kSegs = 64;

__device__ inline CounterStripe* stripe_of(CounterStripe* s, uint32_t wave) { return s + (wave & (kStripes - 1)); }

// Output fills of one call (single words: their waves reserve whole batches).
struct Fills {
  uint32_t p1, p2, c, r;
  uint32_t g;  // two-pass narrow phase: surface pairs through the SAT gate
};

}  // namespace kdop_amd
