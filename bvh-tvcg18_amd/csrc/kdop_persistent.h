// kdop_persistent.h -- persistent tandem traversal.
//
// Three launches per call instead of one per BVTT level:
//   k_seed  (one block per body pair)  root tests + top-level chunk-root pairs
//   k_ramp  (one 512-thread block)     breadth-first levels, synchronised with
//                                      __syncthreads, while the frontier is
//                                      narrower than the persistent grid needs
//   k_persist (persistent grid)        waves take guided chunks of that
//                                      frontier (one atomic per chunk) and
//                                      expand them depth first from a stack
//                                      in LDS; a full stack spills to a
//                                      global buffer that the next round
//                                      (same kernel) takes as its frontier.
// Agent-scope atomics and polling loads on one word are slow on a multi-XCD
// part (they bypass the per-XCD L2s), so nothing here polls shared words:
// the only shared RMWs are the chunk counter, the output counters (one per
// wave and batch) and the spill counter.
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

constexpr int kMaxRounds = 8;

struct PersistState {
  unsigned int level_fill[64];      // ramp: items per level (level 0 = seeds)
  unsigned int final_level;         // ramp: level whose buffer holds the frontier
  unsigned int final_count;
  unsigned int ramp_done;           // ramp: 1 when the frontier is complete (empty)
  unsigned int next[kMaxRounds];    // persist: chunk counter per round
  unsigned int count[kMaxRounds];   // persist: frontier size per round (round r > 0: spill of round r-1)
  unsigned int overflow;
  unsigned int pad;
};

// Breadth-first levels in one workgroup until the frontier reaches `target`
// items, runs empty, or `max_levels` levels are done.  Buffers ping-pong:
// level l is read from buf[l & 1] and written to buf[(l + 1) & 1].
template <int K>
__global__ __launch_bounds__(512) void k_ramp(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                               PersistState* __restrict__ S, uint32_t target, uint32_t max_levels) {
  const int lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  for (uint32_t l = 0;; ++l) {
    const uint32_t n = min(__hip_atomic_load(&S->level_fill[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), cap);
    if (n == 0 || n >= target || l >= max_levels || l >= 62) {
      if (threadIdx.x == 0) {
        S->final_level = l;
        S->final_count = n;
        S->count[0] = n;
        S->ramp_done = (n == 0) ? 1u : 0u;
      }
      break;
    }
    const Item* in = (l & 1) ? buf1 : buf0;
    Item* out = (l & 1) ? buf0 : buf1;
    for (uint32_t base = wave * 64; base < n; base += nwaves * 64) {
      const uint32_t i = base + lane;
      Expansion x;
      x.push = x.leaf = x.surf = 0;
      x.tests = x.descent = x.broot = 0;
      Item it{0, 0, 0};
      if (i < n) {
        it = in[i];
        expand_item<K>(T.nodes, it.a, it.b, x);
      }
      tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
      const uint32_t o = wave_reserve(&S->level_fill[l + 1], __popc(x.push));
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((x.push >> k) & 1u) {
          const uint32_t slot = o + __popc(x.push & ((1u << k) - 1u));
          if (slot < cap) out[slot] = Item{x.ca[k >> 1], x.cb[k & 1], it.pair};
          else atomicOr(&S->overflow, 1u);
        }
      emit_leaves(T, it.pair, x);
    }
    __syncthreads();
  }
  wave_add64(&T.cnt->tests, tests);
  wave_add64(&T.cnt->descent, descent);
  wave_add64(&T.cnt->branch_overlaps, broot);
  wave_add64(&T.cnt->p1, p1s);
}

// One round of the persistent traversal over frontier `in` (S->count[round]
// items); children that do not fit the LDS stack go to `spill`, which becomes
// the frontier of round + 1.
template <int K, int LS>
__global__ __launch_bounds__(64) void k_persist(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                PersistState* __restrict__ S, uint32_t round) {
  static_assert((LS & (LS - 1)) == 0 && LS >= 256, "LDS stack: power of two, >= 256");
  __shared__ uint32_t sa[LS], sb[LS], sp[LS];
  const int lane = lane_id();
  // the ramp left its frontier in buf[final_level & 1]; each round's spill
  // goes to the other buffer and is the next round's frontier
  const bool odd = ((S->final_level + round) & 1u) != 0u;
  const Item* __restrict__ in = odd ? buf1 : buf0;
  Item* __restrict__ spill = odd ? buf0 : buf1;
  const uint32_t n_front = min(S->count[round], cap);
  const uint32_t nwaves = gridDim.x;
  uint32_t top = 0;
  uint32_t next_est = 0;  // last chunk start seen, for the guided chunk size
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  for (;;) {
    Item it{0, 0, 0};
    uint32_t n = 0;
    if (top == 0) {
      // guided self-scheduling: chunks shrink as the frontier runs out
      const uint32_t remaining = n_front > next_est ? n_front - next_est : 0u;
      uint32_t want = remaining / (2u * nwaves);
      want = want < 1u ? 1u : (want > 64u ? 64u : want);
      uint32_t start = 0;
      if (lane == 0) start = atomicAdd(&S->next[round], want);
      start = __shfl(start, 0, 64);
      if (start >= n_front) break;
      next_est = start + want;
      n = min(want, n_front - start);
      if ((uint32_t)lane < n) it = in[start + lane];
    } else {
      n = min(64u, top);
      if ((uint32_t)lane < n) {
        const uint32_t s = top - n + lane;
        it = Item{sa[s], sb[s], sp[s]};
      }
      top -= n;
    }
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    if ((uint32_t)lane < n) expand_item<K>(T.nodes, it.a, it.b, x);
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    const uint32_t m = __popc(x.push);
    const uint32_t incl = wave_incl_scan(m);
    const uint32_t total = __shfl(incl, 63, 64);
    if (total) {
      if (top + total <= (uint32_t)LS) {
        uint32_t o = top + incl - m;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((x.push >> k) & 1u) {
            sa[o] = x.ca[k >> 1]; sb[o] = x.cb[k & 1]; sp[o] = it.pair;
            ++o;
          }
        top += total;
      } else {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&S->count[round + 1], total);
        base = __shfl(base, 0, 64);
        uint32_t o = base + incl - m;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((x.push >> k) & 1u) {
            if (o < cap) spill[o] = Item{x.ca[k >> 1], x.cb[k & 1], it.pair};
            else atomicOr(&S->overflow, 1u);
            ++o;
          }
      }
    }
    __syncthreads();  // one-wave block: stack writes land before the next pop
    emit_leaves(T, it.pair, x);
  }
  wave_add64(&T.cnt->tests, tests);
  wave_add64(&T.cnt->descent, descent);
  wave_add64(&T.cnt->branch_overlaps, broot);
  wave_add64(&T.cnt->p1, p1s);
}

}  // namespace kdop_amd
