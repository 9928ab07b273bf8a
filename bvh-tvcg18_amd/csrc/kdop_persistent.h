// kdop_persistent.h -- persistent tandem traversal: one launch, per-wave work
// stacks in LDS, a global ring of work items for seeding and work stealing.
//
// Every wave loops: take up to 64 node pairs (from the top of its own LDS
// stack, else a batch from the global ring), expand them with the same rule
// as the level-synchronous kernel (expand_item), push the overlapping child
// pairs with one wave-wide prefix sum (to its stack, or to the ring when the
// stack is full), write leaf pairs with one atomic per wave, and donate the
// bottom of a deep stack to the ring when the ring runs low.  Termination: a
// global count of items pushed but not yet expanded; a wave with no work
// exits once it reads zero.  Every wait is bounded; running out sets
// QueueState::error and the host re-runs the call level by level.
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

constexpr unsigned long long kFreeSlot = ~0ull;

__device__ inline unsigned long long ld_relaxed(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_relaxed_i(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

struct PersistArgs {
  const float* nodes;
  DevCounters* cnt;
  QueueState* qs;
  unsigned long long* ring;
  unsigned long long qmask;          // ring capacity - 1 (power of two)
  LeafPair* p1;                      // null when P1 is not kept
  LeafPair* p2;
  uint32_t* p1_fill;
  uint32_t* p2_fill;
  uint32_t p1_cap, p2_cap;
  const unsigned long long* pair_keys;  // sorted (body_a << 32 | body_b)
  const uint32_t* pair_index;           // call-order pair index per sorted key
  uint32_t npairs;
  uint32_t donate_min;   // donate when the local stack holds more than this
  uint32_t ring_low;     // ... and the ring holds fewer items than this
  uint32_t spin_limit;   // bound of every wait loop (iterations)
};

__device__ inline uint32_t pair_of(const PersistArgs& P, uint32_t body_a, uint32_t body_b) {
  if (P.npairs <= 1) return 0;
  const unsigned long long key = ((unsigned long long)body_a << 32) | body_b;
  uint32_t lo = 0, hi = P.npairs;
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (P.pair_keys[m] <= key) lo = m; else hi = m;
  }
  return P.pair_index[lo];
}

// Seed the ring: one 64-lane block per body pair; root test, then all pairs
// of top-level chunk roots (kdop_tandem_traversal.h:164-195).
template <int K>
__global__ __launch_bounds__(64) void k_seed_ring(TravArgs T, PersistArgs P) {
  const uint32_t pair = blockIdx.x;
  const uint2 bp = T.pairs[pair];
  const BodyDev& BA = T.bodies[bp.x];
  const BodyDev& BB = T.bodies[bp.y];
  NodeRec<K> ra, rb;
  load_rec<K>(T.nodes, BA.node_base, ra);
  load_rec<K>(T.nodes, BB.node_base, rb);
  const bool ov = dop_overlap<K>(ra.v, rb.v);
  if (lane_id() == 0) {
    atomicAdd(&T.cnt->root_tests, 1ull);
    atomicAdd(&T.cnt->tests, 1ull);
    if (ov) atomicAdd(&T.cnt->branch_pairs, (unsigned long long)BA.nbranch * BB.nbranch);
  }
  if (!ov) return;
  const uint32_t combos = BA.ntop * BB.ntop;
  unsigned long long tests = 0, broot = 0, p1s = 0;
  for (uint32_t base = 0; base < combos; base += 64) {
    const uint32_t w = base + lane_id();
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.ca[0] = x.cb[0] = 0;
    if (w < combos) {
      const uint32_t i = w / BB.ntop, j = w - i * BB.ntop;
      x.ca[0] = T.tops[BA.top_base + i];
      x.cb[0] = T.tops[BB.top_base + j];
      NodeRec<K> X, Y;
      load_rec<K>(T.nodes, x.ca[0], X);
      load_rec<K>(T.nodes, x.cb[0], Y);
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
        }
      }
    }
    const uint32_t incl = wave_incl_scan(x.push);
    const uint32_t total = __shfl(incl, 63, 64);
    unsigned long long slot0 = 0;
    if (lane_id() == 0 && total) {
      atomicAdd(&P.qs->pending, (int)total);
      slot0 = atomicAdd(&P.qs->tail, (unsigned long long)total);
    }
    slot0 = __shfl(slot0, 0, 64);
    if (x.push) P.ring[(slot0 + incl - 1) & P.qmask] = ((unsigned long long)x.ca[0] << 32) | x.cb[0];
    p1s += x.leaf;
    emit_leaves(T, pair, x);
  }
  wave_add64(&T.cnt->tests, tests);
  wave_add64(&T.cnt->branch_overlaps, broot);
  wave_add64(&T.cnt->p1, p1s);
}

// Write this lane's leaf pairs; the body pair comes from the leaf records.
__device__ inline void emit_leaves_p(const PersistArgs& P, const Expansion& x, uint32_t body_a, uint32_t body_b) {
  uint32_t pair = 0;
  if (x.leaf) pair = pair_of(P, body_a, body_b);
  if (P.p1) {
    const uint32_t o1 = wave_reserve(P.p1_fill, __popc(x.leaf));
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if ((x.leaf >> k) & 1u) {
        const uint32_t o = o1 + __popc(x.leaf & ((1u << k) - 1u));
        if (o < P.p1_cap) P.p1[o] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
      }
  }
  const uint32_t o2 = wave_reserve(P.p2_fill, __popc(x.surf));
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((x.surf >> k) & 1u) {
      const uint32_t o = o2 + __popc(x.surf & ((1u << k) - 1u));
      if (o < P.p2_cap) P.p2[o] = LeafPair{pair, x.ta[k >> 1], x.tb[k & 1]};
    }
}

// Push `cnt` items of this lane (item k of `items`) to the ring.  Every lane
// of the wave must call it.  Waits (bounded) for a slot still being read.
__device__ inline bool ring_push(const PersistArgs& P, const unsigned long long* items, uint32_t cnt, uint32_t total) {
  unsigned long long base = 0;
  if (lane_id() == 0 && total) base = atomicAdd(&P.qs->tail, (unsigned long long)total);
  base = __shfl(base, 0, 64);
  const uint32_t incl = wave_incl_scan(cnt);
  bool ok = true;
  for (uint32_t k = 0; k < cnt; ++k) {
    unsigned long long* slot = P.ring + ((base + incl - cnt + k) & P.qmask);
    uint32_t spins = 0;
    while (ld_relaxed(slot) != kFreeSlot) {
      if (++spins > P.spin_limit) { ok = false; break; }
      __builtin_amdgcn_s_sleep(1);
    }
    if (!ok) break;
    st_relaxed(slot, items[k]);
  }
  return ok;
}

template <int K, int LS>
__global__ __launch_bounds__(64) void k_traverse_persistent(PersistArgs P) {
  static_assert((LS & (LS - 1)) == 0, "LDS stack size must be a power of two");
  __shared__ unsigned long long stack[LS];
  const int lane = lane_id();
  uint32_t top = 0, bot = 0;  // stack holds items [bot, top) modulo LS (wave-uniform)
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0, steals = 0, donations = 0;
  uint32_t idle = 0;
  bool failed = false;
  for (;;) {
    // ---- take up to 64 items
    unsigned long long item = kFreeSlot;
    uint32_t n = 0;
    if (top != bot) {
      n = min(64u, top - bot);
      if ((uint32_t)lane < n) item = stack[(top - n + lane) & (LS - 1)];
      top -= n;
    } else {
      unsigned long long h = 0;
      if (lane == 0) {
        for (uint32_t tries = 0;; ++tries) {
          const unsigned long long hh = ld_relaxed(&P.qs->head);
          const unsigned long long tt = ld_relaxed(&P.qs->tail);
          if (hh >= tt || tries > 4096) { n = 0; break; }  // contended: retry after the idle check
          const unsigned long long avail = tt - hh;
          n = (uint32_t)(avail < 64 ? avail : 64);
          if (atomicCAS(&P.qs->head, hh, hh + n) == hh) { h = hh; break; }
        }
      }
      n = __shfl(n, 0, 64);
      h = __shfl(h, 0, 64);
      if (n == 0) {
        int pend = 0;
        if (lane == 0) pend = ld_relaxed_i(&P.qs->pending);
        pend = __shfl(pend, 0, 64);
        if (pend == 0) break;  // every pushed item has been expanded: done
        if (++idle > P.spin_limit) { failed = true; break; }
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      ++steals;
      if ((uint32_t)lane < n) {
        unsigned long long* slot = P.ring + ((h + lane) & P.qmask);
        uint32_t spins = 0;
        for (;;) {
          item = ld_relaxed(slot);
          if (item != kFreeSlot) break;
          if (++spins > P.spin_limit) { failed = true; break; }
          __builtin_amdgcn_s_sleep(1);
        }
        st_relaxed(slot, kFreeSlot);
      }
      if (__any(failed)) { failed = true; break; }
    }
    idle = 0;
    // ---- expand
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    uint32_t body_a = 0, body_b = 0;
    if ((uint32_t)lane < n) {
      expand_item<K>(P.nodes, (uint32_t)(item >> 32), (uint32_t)item, x);
      body_a = x.body_a;
      body_b = x.body_b;
    }
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    // ---- push children: local stack when it has room, else the ring
    const uint32_t m = __popc(x.push);
    const uint32_t incl = wave_incl_scan(m);
    const uint32_t total = __shfl(incl, 63, 64);
    if (total && lane == 0) atomicAdd(&P.qs->pending, (int)total);  // before they become visible
    unsigned long long kids[4];
    uint32_t nk = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if ((x.push >> k) & 1u) kids[nk++] = ((unsigned long long)x.ca[k >> 1] << 32) | x.cb[k & 1];
    if (total) {
      if ((top - bot) + total <= (uint32_t)LS) {
        uint32_t o = top + incl - m;
        for (uint32_t k = 0; k < nk; ++k) stack[(o + k) & (LS - 1)] = kids[k];
        top += total;
        __syncthreads();  // one-wave block: orders the stack writes before the next pop
      } else {
        ++donations;
        if (!ring_push(P, kids, nk, total)) failed = true;
        if (__any(failed)) { failed = true; }
      }
    }
    emit_leaves_p(P, x, body_a, body_b);
    // ---- donate the bottom of a deep stack while the ring runs low
    if (!failed && (top - bot) > P.donate_min) {
      unsigned long long occ = 0;
      if (lane == 0) occ = ld_relaxed(&P.qs->tail) - ld_relaxed(&P.qs->head);
      occ = __shfl(occ, 0, 64);
      if (occ < P.ring_low) {
        const unsigned long long it = stack[(bot + lane) & (LS - 1)];
        bot += 64;
        ++donations;
        if (!ring_push(P, &it, 1u, 64u)) failed = true;
      }
    }
    if (lane == 0 && n) atomicSub(&P.qs->pending, (int)n);
    if (__any(failed)) { failed = true; break; }
  }
  if (failed && lane == 0) atomicOr(&P.qs->error, 1u);
  wave_add64(&P.cnt->tests, tests);
  wave_add64(&P.cnt->descent, descent);
  wave_add64(&P.cnt->branch_overlaps, broot);
  wave_add64(&P.cnt->p1, p1s);
  if (lane == 0) {
    atomicAdd(&P.qs->steals, steals);
    atomicAdd(&P.qs->donations, donations);
  }
}

}  // namespace kdop_amd
