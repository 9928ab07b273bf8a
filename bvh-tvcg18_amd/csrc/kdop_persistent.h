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
  unsigned int max_iters;            // persist: most loop iterations of one wave (all rounds)
  unsigned long long iters;          // persist: loop iterations summed over waves
  unsigned long long lanes;          // persist: items expanded (active lanes) summed over waves
  unsigned long long grabs;          // persist: chunks taken from the frontier
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
      Item it{0, 0, 0, 0, 0};
      if (i < n) {
        it = in[i];
        expand_item<K>(T.nodes, it, x);
      }
      tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
      const uint32_t o = wave_reserve(&S->level_fill[l + 1], __popc(x.push));
      emit_items(x, out, o, cap, &S->overflow);
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
// items).  Each iteration a wave expands up to 64 pairs: the top of its LDS
// stack, topped up with a guided chunk of the frontier while it lasts.
// Children go on the stack, or to `spill` (the next round's frontier) when it
// is full.  Once the frontier is gone, a wave keeps at most `keep` items and
// spills the bottom (oldest, largest subtrees) of a deeper stack, so the next
// round spreads the tail over every wave.
template <int K, int LS>
__global__ __launch_bounds__(64) void k_persist(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                PersistState* __restrict__ S, uint32_t round, uint32_t keep) {
  static_assert((LS & (LS - 1)) == 0 && LS >= 256, "LDS stack: power of two, >= 256");
  __shared__ uint32_t sa0[LS], sa1[LS], sb0[LS], sb1[LS], sp[LS];
  const int lane = lane_id();
  // the ramp left its frontier in buf[final_level & 1]; each round's spill
  // goes to the other buffer and is the next round's frontier
  const bool odd = ((S->final_level + round) & 1u) != 0u;
  const Item* __restrict__ in = odd ? buf1 : buf0;
  Item* __restrict__ spill = odd ? buf0 : buf1;
  const uint32_t n_front = min(S->count[round], cap);
  const uint32_t nwaves = gridDim.x;
  uint32_t top = 0, bot = 0;  // stack = [bot, top) modulo LS, wave-uniform
  uint32_t next_est = 0;      // last chunk start seen, for the guided chunk size
  bool front_done = n_front == 0;
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  uint32_t iters = 0, lanes = 0, grabs = 0;
  for (;;) {
    // ---- batch: top of the stack, then a chunk of the frontier
    const uint32_t ns = min(64u, top - bot);
    Item it{0, 0, 0, 0, 0};
    if ((uint32_t)lane < ns) {
      const uint32_t s = (top - ns + lane) & (LS - 1);
      it = Item{sa0[s], sa1[s], sb0[s], sb1[s], sp[s]};
    }
    top -= ns;
    uint32_t ng = 0;
    if (ns < 64 && !front_done) {
      const uint32_t remaining = n_front > next_est ? n_front - next_est : 0u;
      uint32_t want = remaining / (2u * nwaves);
      want = want < 1u ? 1u : (want > 64u - ns ? 64u - ns : want);
      uint32_t start = 0;
      if (lane == 0) start = atomicAdd(&S->next[round], want);
      start = __shfl(start, 0, 64);
      if (start >= n_front) {
        front_done = true;
      } else {
        ++grabs;
        next_est = start + want;
        ng = min(want, n_front - start);
        if ((uint32_t)lane >= ns && (uint32_t)lane < ns + ng) it = in[start + lane - ns];
      }
    } else if (!front_done && (iters & 15u) == 15u) {
      // a wave living off its own stack looks at the frontier now and then
      uint32_t nx = 0;
      if (lane == 0) nx = __hip_atomic_load(&S->next[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      front_done = __shfl(nx, 0, 64) >= n_front;
    }
    const uint32_t n = ns + ng;
    if (n == 0) break;
    // ---- expand
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    if ((uint32_t)lane < n) expand_item<K>(T.nodes, it, x);
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    ++iters;
    lanes += n;
    // ---- children: stack when they fit, else the next round
    const uint32_t nk = __popc(x.push);
    const uint32_t incl = wave_incl_scan(nk);
    const uint32_t total = __shfl(incl, 63, 64);
    if (total) {
      if ((top - bot) + total <= (uint32_t)LS) {
        uint32_t o = top + incl - nk;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((x.push >> k) & 1u) {
            const uint32_t s2 = (o++) & (LS - 1);
            const Item& t = x.kids[k];
            sa0[s2] = t.a0; sa1[s2] = t.a1; sb0[s2] = t.b0; sb1[s2] = t.b1; sp[s2] = t.pair;
          }
        top += total;
      } else {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&S->count[round + 1], total);
        base = __shfl(base, 0, 64);
        emit_items(x, spill, base + incl - nk, cap, &S->overflow);
      }
    }
    // ---- tail: hand the bottom of a deep stack to the next round
    if (front_done && (top - bot) > keep + 64) {
      const uint32_t s3 = (bot + lane) & (LS - 1);
      const Item mine = Item{sa0[s3], sa1[s3], sb0[s3], sb1[s3], sp[s3]};
      bot += 64;
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&S->count[round + 1], 64u);
      base = __shfl(base, 0, 64) + lane;
      if (base < cap) spill[base] = mine;
      else atomicOr(&S->overflow, 1u);
    }
    __syncthreads();  // one-wave block: stack writes land before the next pop
    emit_leaves(T, it.pair, x);
  }
  if (lane == 0) {
    atomicMax(&S->max_iters, iters);
    atomicAdd(&S->iters, (unsigned long long)iters);
    atomicAdd(&S->lanes, (unsigned long long)lanes);
    atomicAdd(&S->grabs, (unsigned long long)grabs);
  }
  wave_add64(&T.cnt->tests, tests);
  wave_add64(&T.cnt->descent, descent);
  wave_add64(&T.cnt->branch_overlaps, broot);
  wave_add64(&T.cnt->p1, p1s);
}

}  // namespace kdop_amd
