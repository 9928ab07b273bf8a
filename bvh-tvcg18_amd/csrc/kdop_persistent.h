// kdop_persistent.h -- persistent tandem traversal.
//
// A few launches per call instead of one per BVTT level:
//   k_seed  (grid over seed pairs)     root tests + chunk-root pairs
//   k_bfs_level (grid, a few launches) breadth-first levels while the frontier
//                                      is narrower than the persistent grid
//                                      needs (decided on the device)
//   k_persist (persistent grid)        waves take guided chunks of that
//                                      frontier and expand them depth first
//                                      from a stack in LDS; a full stack
//                                      spills to a global buffer that the
//                                      next round (same kernel) takes as its
//                                      frontier.
// Agent-scope atomics on one word are slow on a multi-XCD part: they are
// performed beyond the per-XCD L2s and serialize, so a counter that a
// thousand waves hit costs more than the traversal.  Every hot counter is
// therefore spread: each round's frontier is cut into kSegs segments, each
// with its own chunk and spill counters on a line of its own; statistics go
// to striped counters; leaf pairs are staged in LDS and written in batches.
#pragma once

#include "kdop_kernels.h"

namespace kdop_amd {

constexpr int kMaxRounds = 8;

// Counters of one segment of one round's frontier.
struct alignas(256) SegCtr {
  uint32_t fill;  // items spilled into the segment by the previous round
  uint32_t next;  // chunk counter of the round that reads the segment
};

// Summary of one call, written by k_sum_counters into pinned host memory.
struct CallSummary {
  unsigned long long total[kStripeFields];  // CounterStripe fields summed (max_iters: max)
  uint32_t overflow;
  Fills fills;
  uint32_t level_fill[64];
  uint32_t final_level, final_count;
  uint32_t round_total[kMaxRounds];   // items in round r's frontier
  uint32_t round_maxseg[kMaxRounds];  // largest segment fill of round r (r >= 1)
  uint32_t reset;                      // 1: the fold zeroed the call state
  uint32_t final_set;                  // a breadth-first launch recorded the frontier
  uint32_t ready;                      // set to 1 last (system-scope release); the host clears it
};

// Device state of one traversal call, zeroed by the host before the call.
struct CallState {
  CounterStripe stripe[kStripes];
  SegCtr seg[kMaxRounds][kSegs];
  alignas(256) unsigned long long done[kMaxRounds];  // per round: segments whose last chunk was taken
  alignas(256) unsigned int overflow;                // bit 0 frontier, 1 P1, 2 P2, 3 contacts
  alignas(256) Fills fills;
  alignas(256) uint32_t level_fill[64];     // items per BFS level (level 0 = seeds)
  uint32_t final_level, final_count, final_set;  // level whose buffer holds the frontier, its size
  SegCtr lseg[3][kSegs];  // segmented levels: per-segment fills, level l in set l % 3
};

// One breadth-first level of the BVTT over the whole grid: expand every item
// of level `level` (buf[level & 1]) into level + 1.  The launch that finds the
// level empty, at least `target` items wide, or equal to `last` expands
// nothing and records it as the frontier (final_level / final_count); later
// launches of the same call return at once.  The host queues `last` + 1 of
// them without reading any size back.  With target = ~0u and last = BVTT
// depth this is the whole level-synchronous traversal.
//
// A level is one contiguous run (nseg 1: its size in level_fill[level]) or
// kSegs segments of cap / kSegs slots, each filled through a counter of its
// own (lseg[level % 3][s]; a block writes segment blockIdx % kSegs).  One
// counter for a whole wide level is hit once per block iteration from every
// XCD, and those atomics serialize: at tens of thousands per level they, not
// the tests, set the level's time.  Item i of a segmented level is found from
// the prefix sums of the segment fills.  Level l clears set (l + 2) % 3, the
// previous level's input, for level l + 1 to write.
//
// QUAD: four lanes per item (expand_quad), so a block expands 64 items per
// pass with 8 record loads and 32 slab compares per lane instead of 256 items
// with 32 loads and up to 104 compares per lane.  The narrow levels of the
// persistent strategy's ramp then spread over four times as many waves and
// each lane's chain of loads and tests is a quarter as long.
template <int K, bool QUAD>
__global__ __launch_bounds__(256) void k_bfs_level(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                   CallState* __restrict__ C, uint32_t level, uint32_t target, uint32_t last,
                                                   uint32_t nseg_in, uint32_t nseg_out) {
  __shared__ uint32_t red[5];
  __shared__ uint32_t pre[kSegs + 1];
  __shared__ uint32_t raw_total;
  __shared__ P2Stage st2;
  __shared__ uint32_t stopped;
  // Wave 0 reads the stop flag and the level's size (its segment fills) at
  // once, one round trip, and shares them through LDS.  One read per block:
  // block 0 of the launch that stops the levels sets final_set while other
  // blocks of that launch start, and every wave of a block must take the same
  // branch.  The sizes are used only when the flag is clear.
  const uint32_t segcap_in = cap / nseg_in;
  // a contiguous level's first item of this thread is loaded before the
  // level's size is known (its slot exists whatever the size; the value is
  // used only when the slot holds an item), so that load overlaps the size read
  const Item* __restrict__ in = (level & 1) ? buf1 : buf0;
  constexpr uint32_t LPI = QUAD ? 4u : 1u;  // lanes per item
  const uint32_t i_first = (blockIdx.x * blockDim.x + threadIdx.x) / LPI;
  const bool early = nseg_in == 1u && i_first < cap;
  Item it_first{0, 0, 0, 0, 0};
  if (early) it_first = in[i_first];
  if (threadIdx.x < 64u) {
    const uint32_t lane = threadIdx.x;
    const uint32_t f = nseg_in == 1u ? (lane == 0u ? C->level_fill[level] : 0u) : C->lseg[level % 3u][lane].fill;
    if (lane == 0u) stopped = __hip_atomic_load(&C->final_set, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nseg_in == 1u) {
      if (lane == 0u) { pre[kSegs] = min(f, cap); raw_total = f; }
    } else {  // prefix sums of the segment fills
      uint32_t x = min(f, segcap_in), t = f;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
        t += __shfl_xor(t, d, 64);
      }
      pre[lane + 1u] = x;
      if (lane == 0u) { pre[0] = 0u; raw_total = t; }
    }
  }
  __syncthreads();
  if (stopped) return;
  const uint32_t n = pre[kSegs];
  if (n == 0 || n >= target || level >= last) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      C->final_level = level;
      C->final_count = n;
      C->final_set = 1u;
    }
    return;
  }
  if (blockIdx.x == 0) {
    if (nseg_in > 1u && threadIdx.x == 0) C->level_fill[level] = raw_total;  // the level's size (all that was reserved)
    if (nseg_out > 1u && threadIdx.x < (uint32_t)kSegs) C->lseg[(level + 2u) % 3u][threadIdx.x].fill = 0u;
  }
  Item* __restrict__ out = (level & 1) ? buf0 : buf1;
  const uint32_t so = blockIdx.x % nseg_out, segcap_out = cap / nseg_out;
  uint32_t* out_fill = nseg_out == 1u ? &C->level_fill[level + 1] : &C->lseg[(level + 1u) % 3u][so].fill;
  Item* __restrict__ out_seg = out + (size_t)so * segcap_out;
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  if (threadIdx.x == 0) st2.n = 0u;
  __syncthreads();
  if constexpr (QUAD) {
    const float4* __restrict__ qbase = quad_lane_base<K>(T.nodes, threadIdx.x & 3u);
    const uint32_t ipb = blockDim.x / 4u;  // items per block pass
    for (uint32_t base = blockIdx.x * ipb; base < n; base += gridDim.x * ipb) {
      const uint32_t i = base + (threadIdx.x >> 2);
      const bool live = i < n;
      Item it{0, 0, 0, 0, 0};
      if (live) {
        if (early && i == i_first) {
          it = it_first;
        } else {
          uint32_t idx = i;
          if (nseg_in > 1u) {  // segment s with pre[s] <= i < pre[s + 1]
            uint32_t lo = 0, hi = kSegs;
            while (hi - lo > 1u) {
              const uint32_t mid = (lo + hi) >> 1;
              if (pre[mid] <= i) lo = mid; else hi = mid;
            }
            idx = lo * segcap_in + (i - pre[lo]);
          }
          it = in[idx];
        }
      }
      QuadRec<K> A0, A1, B0, B1;
      QuadOut o;
      expand_quad<K>(qbase, it, live, A0, A1, B0, B1, o);
      tests += o.test; descent += o.descent; broot += o.broot; p1s += o.leaf;
      const uint32_t slot = block_reserve(out_fill, o.push ? 1u : 0u, red);
      if (o.push) {
        if (slot < segcap_out) out_seg[slot] = o.kid;
        else atomicOr(T.overflow, 1u);
      }
      if (T.p1) {
        const uint32_t o1 = wave_reserve(T.p1_fill, o.leaf ? 1u : 0u);
        if (o.leaf) {
          if (o1 < T.p1_cap) T.p1[o1] = LeafPair{it.pair, o.ta, o.tb};
          else atomicOr(T.overflow, 2u);
        }
      }
      stage_p2_lane(st2, it.pair, o.surf, o.ta, o.tb);
      flush_p2(T, st2, false);
    }
  } else {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    Item it{0, 0, 0, 0, 0};
    if (i < n) {
      if (early && i == i_first) {
        it = it_first;
      } else {
        uint32_t idx = i;
        if (nseg_in > 1u) {  // segment s with pre[s] <= i < pre[s + 1]
          uint32_t lo = 0, hi = kSegs;
          while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= i) lo = mid; else hi = mid;
          }
          idx = lo * segcap_in + (i - pre[lo]);
        }
        it = in[idx];
      }
      expand_item<K>(T.nodes, it, x);
    }
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    const uint32_t o = block_reserve(out_fill, __popc(x.push), red);
    emit_items(x, out_seg, o, segcap_out, T.overflow);
    emit_p1(T, it.pair, x);
    stage_p2(st2, it.pair, x);
    flush_p2(T, st2, false);
  }
  }
  flush_p2(T, st2, true);
  CounterStripe* st = stripe_of(T.stripes, wave_global());
  wave_add64(&st->tests, tests);
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
}

// First segment after `cur` (cyclically) whose bit in `known` is clear;
// `known` must have a clear bit.
__device__ inline uint32_t next_open(unsigned long long known, uint32_t cur) {
  const unsigned long long open = ~known;
  const unsigned long long above = cur >= 63u ? 0ull : (open & (~0ull << (cur + 1u)));
  return (uint32_t)(__ffsll((unsigned long long)(above ? above : open)) - 1);
}

// Write the staged leaf pairs of one wave: all of them to P1 (when kept),
// the surface subset (bit 31 of the pair word) to P2.  One atomic per list.
__device__ inline void flush_leaves(const TravArgs& T, const uint32_t* lp_pair, const uint32_t* lp_ta, const uint32_t* lp_tb,
                                    uint32_t nst) {
  const uint32_t lane = lane_id();
  if (T.p1) {
    uint32_t b = 0;
    if (lane == 0) b = atomicAdd(T.p1_fill, nst);
    b = lane0(b);
    for (uint32_t i = lane; i < nst; i += 64) {
      const uint32_t o = b + i;
      if (o < T.p1_cap) T.p1[o] = LeafPair{lp_pair[i] & 0x7FFFFFFFu, lp_ta[i], lp_tb[i]};
      else atomicOr(T.overflow, 2u);
    }
  }
  uint32_t n2 = 0;
  for (uint32_t i0 = 0; i0 < nst; i0 += 64) {
    const uint32_t i = i0 + lane;
    n2 += (uint32_t)__popcll(__ballot(i < nst && (lp_pair[i] >> 31)));
  }
  if (n2 == 0) return;
  uint32_t b2 = 0;
  if (lane == 0) b2 = atomicAdd(T.p2_fill, n2);
  b2 = lane0(b2);
  const unsigned long long below = (1ull << lane) - 1ull;
  for (uint32_t i0 = 0; i0 < nst; i0 += 64) {
    const uint32_t i = i0 + lane;
    const bool s = i < nst && (lp_pair[i] >> 31);
    const unsigned long long m = __ballot(s);
    if (s) {
      const uint32_t o = b2 + (uint32_t)__popcll(m & below);
      if (o < T.p2_cap) T.p2[o] = LeafPair{lp_pair[i] & 0x7FFFFFFFu, lp_ta[i], lp_tb[i]};
      else atomicOr(T.overflow, 4u);
    }
    b2 += (uint32_t)__popcll(m);
  }
}

// One round of the persistent traversal over this round's frontier (round 0:
// the last breadth-first level, contiguous; round r > 0: the segments that round
// r - 1 spilled).  Each iteration a wave expands up to 64 pairs: the top of
// its LDS stack, topped up with a guided chunk of its current segment while
// the frontier lasts.  Children go on the stack, or to the next round's
// segments when it is full.  Once the frontier is gone, a wave keeps at most
// `keep` items and spills the bottom (oldest, largest subtrees) of a deeper
// stack, so the next round spreads the tail over every wave.
template <int K, int LS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_persist(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                CallState* __restrict__ C, uint32_t round, uint32_t keep, uint32_t last) {
  static_assert((LS & (LS - 1)) == 0 && LS >= 256, "LDS stack: power of two, >= 256");
  static_assert(kSegs == 64, "one lane per segment");
  constexpr uint32_t SB = 512;  // staged leaf pairs; flushed above SB - 256 (one iteration adds <= 256)
  __shared__ uint32_t sa0[LS], sa1[LS], sb0[LS], sb1[LS], sp[LS];
  __shared__ uint32_t lp_pair[SB], lp_ta[SB], lp_tb[SB];
  const uint32_t lane = lane_id();
  // the breadth-first levels left the frontier in buf[final_level & 1]; each round's spill
  // goes to the other buffer and is the next round's frontier.  A breadth-first launch
  // that stopped early recorded the frontier; when none did, every queued level was
  // expanded and round 0 takes level `last` (every block derives it the same way, and
  // block 0 records it for the later rounds and the summary)
  uint32_t final_level = C->final_level, final_count = 0;
  if (round == 0) {
    const bool recorded = C->final_set != 0u;
    if (!recorded) final_level = last;
    final_count = recorded ? C->final_count : min(C->level_fill[last], cap);
    if (!recorded && blockIdx.x == 0 && threadIdx.x == 0) {
      C->final_level = final_level;
      C->final_count = final_count;
    }
  }
  const bool odd = ((final_level + round) & 1u) != 0u;
  const Item* __restrict__ in = odd ? buf1 : buf0;
  Item* __restrict__ spill = odd ? buf0 : buf1;
  const uint32_t segcap = cap / kSegs;
  const uint32_t wave = blockIdx.x;
  const uint32_t wps = max(1u, gridDim.x / (uint32_t)kSegs);  // waves per segment
  // lane s holds segment s of this round's frontier
  uint32_t seg_base, seg_len;
  if (round == 0) {
    const uint32_t n = min(final_count, cap);
    seg_base = (uint32_t)(((unsigned long long)n * lane) / kSegs);
    seg_len = (uint32_t)(((unsigned long long)n * (lane + 1)) / kSegs) - seg_base;
  } else {
    seg_base = lane * segcap;
    seg_len = min(C->seg[round][lane].fill, segcap);
  }
  unsigned long long known = __ballot(seg_len == 0u);  // segments known to be exhausted
  const uint32_t home = wave & (kSegs - 1);
  uint32_t cur = home;
  bool front_done = known == ~0ull;
  if (!front_done && ((known >> cur) & 1ull)) cur = next_open(known, cur);
  uint32_t top = 0, bot = 0;  // stack = [bot, top) modulo LS, wave-uniform
  uint32_t next_est = 0;      // end of the last chunk taken from `cur`
  uint32_t nsp = 0;           // spills so far (rotates the target segment)
  uint32_t nstage = 0;        // staged leaf pairs
  const bool keep_p1 = T.p1 != nullptr;
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  uint32_t iters = 0, lanes = 0, grabs = 0, handed = 0;
  for (;;) {
    top = uniform(top); bot = uniform(bot); cur = uniform(cur); next_est = uniform(next_est);
    // ---- batch: top of the stack, then a chunk of the current segment
    const uint32_t ns = min(64u, top - bot);
    Item it{0, 0, 0, 0, 0};
    if (lane < ns) {
      const uint32_t s = (top - ns + lane) & (LS - 1);
      it = Item{sa0[s], sa1[s], sb0[s], sb1[s], sp[s]};
    }
    top -= ns;
    uint32_t ng = 0;
    bool refresh = false;
    if (ns < 64 && !front_done) {
      const uint32_t len = __builtin_amdgcn_readlane(seg_len, cur), sbase = __builtin_amdgcn_readlane(seg_base, cur);
      const uint32_t remaining = len > next_est ? len - next_est : 0u;
      uint32_t want = remaining / (2u * wps);
      want = want < 1u ? 1u : (want > 64u - ns ? 64u - ns : want);
      uint32_t start = 0;
      if (lane == 0) start = atomicAdd(&C->seg[round][cur].next, want);
      start = lane0(start);
      if (start < len) {
        ++grabs;
        next_est = start + want;
        ng = min(want, len - start);
        if (lane >= ns && lane < ns + ng) it = in[sbase + start + lane - ns];
        if (start + want >= len) {  // this wave took the segment's last chunk
          if (lane == 0) atomicOr(&C->done[round], 1ull << cur);
          known |= 1ull << cur;
        }
      } else {
        known |= 1ull << cur;
        refresh = true;
      }
    } else if (!front_done && (iters & 15u) == 15u) {
      refresh = true;  // a wave living off its own stack looks at the frontier now and then
    }
    if (refresh) {
      unsigned long long d = 0;
      if (lane == 0) d = __hip_atomic_load(&C->done[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      known |= ((unsigned long long)lane0((uint32_t)(d >> 32)) << 32) | lane0((uint32_t)d);
    }
    if (!front_done && ((known >> cur) & 1ull)) {
      if (known == ~0ull) front_done = true;
      else { cur = next_open(known, cur); next_est = 0; }
    }
    const uint32_t n = ns + ng;
    if (n == 0) {
      if (front_done) break;
      continue;  // the current segment ran dry: try the next one
    }
    // ---- expand
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    if (lane < n) expand_item<K>(T.nodes, it, x);
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    ++iters;
    lanes += n;
    // ---- children: stack when they fit, else the next round
    uint32_t total;
    const uint32_t kpre = wave_prefix_small(__popc(x.push), &total);
    if (total) {
      if ((top - bot) + total <= (uint32_t)LS) {
        uint32_t o = top + kpre;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((x.push >> k) & 1u) {
            const uint32_t s2 = (o++) & (LS - 1);
            const Item& t = x.kids[k];
            sa0[s2] = t.a0; sa1[s2] = t.a1; sb0[s2] = t.b0; sb1[s2] = t.b1; sp[s2] = t.pair;
          }
        top += total;
      } else {
        const uint32_t sg = (home + nsp++) & (kSegs - 1);
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, total);
        b = lane0(b);
        emit_items(x, spill + (size_t)sg * segcap, b + kpre, segcap, T.overflow);
      }
    }
    // ---- tail: hand the bottom of a deep stack to the next round
    if (front_done && (top - bot) > keep + 64) {
      const uint32_t s3 = (bot + lane) & (LS - 1);
      const Item mine = Item{sa0[s3], sa1[s3], sb0[s3], sb1[s3], sp[s3]};
      bot += 64;
      const uint32_t sg = (home + nsp++) & (kSegs - 1);
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, 64u);
      b = lane0(b) + lane;
      if (b < segcap) spill[(size_t)sg * segcap + b] = mine;
      else atomicOr(T.overflow, 1u);
      handed += 64u;
    }
    // ---- leaf pairs: stage in LDS (all of P1 when it is kept, else only P2)
    const uint32_t lm = keep_p1 ? x.leaf : x.surf;
    uint32_t ltot;
    const uint32_t lpre = wave_prefix_small(__popc(lm), &ltot);
    if (ltot) {
      uint32_t o = nstage + lpre;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((lm >> k) & 1u) {
          lp_pair[o] = it.pair | (((x.surf >> k) & 1u) << 31);
          lp_ta[o] = x.ta[k >> 1];
          lp_tb[o] = x.tb[k & 1];
          ++o;
        }
      nstage += ltot;
      if (nstage > SB - 256) {
        wave_lds_sync();
        flush_leaves(T, lp_pair, lp_ta, lp_tb, nstage);
        nstage = 0;
      }
    }
    wave_lds_sync();  // stack and staging writes land before the next reads
  }
  wave_lds_sync();
  if (nstage) flush_leaves(T, lp_pair, lp_ta, lp_tb, nstage);
  CounterStripe* st = stripe_of(T.stripes, wave);
  const unsigned long long wtests = wave_sum64(tests);
  if (lane == 0 && wtests) {
    atomicAdd(&st->tests, wtests);
    atomicAdd(&st->ptests, wtests);
  }
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
  if (lane == 0) {
    atomicAdd(&st->iters, (unsigned long long)iters);
    atomicAdd(&st->lanes, (unsigned long long)lanes);
    atomicAdd(&st->grabs, (unsigned long long)grabs);
    if (handed) atomicAdd(&st->handoffs, (unsigned long long)handed);
    atomicMax(&st->max_iters, (unsigned long long)iters);
  }
}

// The same round with W waves per workgroup sharing one LDS stack of W * LS
// items.  Each iteration the block pops up to 64 W items from the top of the
// stack, topped up with a guided chunk of the frontier, expands one per
// thread, and pushes the children at block-wide prefix-sum places, so the
// subtrees of the frontier items the block took are spread over all its
// waves (with one stack per wave, the wave that took the largest subtrees
// runs on alone).  Two barriers per iteration; every value that steers the
// loop is read from LDS after a barrier, so all waves take the same path.
// Wave 0 keeps the frontier state and decides the next grab between the
// barriers.  Spills, the tail hand-off and the leaf staging work as in
// k_persist, per wave.
template <int K, int LS, int W>
__global__ __launch_bounds__(64 * W) void k_persist_blk(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                        CallState* __restrict__ C, uint32_t round, uint32_t keep, uint32_t last) {
  constexpr uint32_t NT = 64u * W;
  constexpr uint32_t S = (uint32_t)LS * W;
  static_assert((S & (S - 1)) == 0 && LS >= 256, "LDS stack: power of two");
  static_assert(kSegs == 64, "one lane of wave 0 per segment");
  constexpr uint32_t SB = 512;  // staged leaf pairs per wave; flushed above SB - 256
  __shared__ uint32_t sa0[S], sa1[S], sb0[S], sb1[S], sp[S];
  __shared__ uint32_t lp_pair[W][SB], lp_ta[W][SB], lp_tb[W][SB];
  __shared__ uint32_t wtot[W];
  __shared__ uint32_t ctl[3];  // next grab: first frontier index, count; frontier done
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  uint32_t final_level = C->final_level, final_count = 0;
  if (round == 0) {
    const bool recorded = C->final_set != 0u;
    if (!recorded) final_level = last;
    final_count = recorded ? C->final_count : min(C->level_fill[last], cap);
    if (!recorded && blockIdx.x == 0 && tid == 0) {
      C->final_level = final_level;
      C->final_count = final_count;
    }
  }
  const bool odd = ((final_level + round) & 1u) != 0u;
  const Item* __restrict__ in = odd ? buf1 : buf0;
  Item* __restrict__ spill = odd ? buf0 : buf1;
  const uint32_t segcap = cap / kSegs;
  const uint32_t bps = max(1u, gridDim.x / (uint32_t)kSegs);  // blocks per segment
  const uint32_t home = blockIdx.x & (kSegs - 1);
  // wave 0 only: lane s holds segment s of this round's frontier
  uint32_t seg_base = 0, seg_len = 0, cur = home, next_est = 0;
  unsigned long long known = ~0ull;
  bool wfront_done = true;  // wave 0's view
  if (w == 0) {
    if (round == 0) {
      const uint32_t n = min(final_count, cap);
      seg_base = (uint32_t)(((unsigned long long)n * lane) / kSegs);
      seg_len = (uint32_t)(((unsigned long long)n * (lane + 1)) / kSegs) - seg_base;
    } else {
      seg_base = lane * segcap;
      seg_len = min(C->seg[round][lane].fill, segcap);
    }
    known = __ballot(seg_len == 0u);
    wfront_done = known == ~0ull;
    if (!wfront_done && ((known >> cur) & 1ull)) cur = next_open(known, cur);
  }
  uint32_t grabs = 0, wave0_iters = 0;
  // wave 0: the grab for an iteration whose stack part is ns items; writes ctl
  auto plan = [&](uint32_t ns) {
    cur = uniform(cur); next_est = uniform(next_est);
    uint32_t gstart = 0, ng = 0;
    bool refresh = false;
    if (ns < NT && !wfront_done) {
      const uint32_t len = __builtin_amdgcn_readlane(seg_len, cur), sbase = __builtin_amdgcn_readlane(seg_base, cur);
      const uint32_t remaining = len > next_est ? len - next_est : 0u;
      uint32_t want = remaining / (2u * bps);
      want = want < 1u ? 1u : (want > NT - ns ? NT - ns : want);
      uint32_t start = 0;
      if (lane == 0) start = atomicAdd(&C->seg[round][cur].next, want);
      start = lane0(start);
      if (start < len) {
        ++grabs;
        next_est = start + want;
        ng = min(want, len - start);
        gstart = sbase + start;
        if (start + want >= len) {  // this block took the segment's last chunk
          if (lane == 0) atomicOr(&C->done[round], 1ull << cur);
          known |= 1ull << cur;
        }
      } else {
        known |= 1ull << cur;
        refresh = true;
      }
    } else if (!wfront_done && (wave0_iters & 15u) == 15u) {
      refresh = true;  // a block living off its stack looks at the frontier now and then
    }
    if (refresh) {
      unsigned long long d = 0;
      if (lane == 0) d = __hip_atomic_load(&C->done[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      known |= ((unsigned long long)lane0((uint32_t)(d >> 32)) << 32) | lane0((uint32_t)d);
    }
    if (!wfront_done && ((known >> cur) & 1ull)) {
      if (known == ~0ull) wfront_done = true;
      else { cur = next_open(known, cur); next_est = 0; }
    }
    if (lane == 0) {
      ctl[0] = gstart;
      ctl[1] = ng;
      ctl[2] = wfront_done ? 1u : 0u;
    }
  };
  if (w == 0) plan(0u);
  __syncthreads();
  uint32_t top = 0, bot = 0;  // stack = [bot, top) modulo S, equal in every thread
  uint32_t nstage = 0;        // this wave's staged leaf pairs
  const bool keep_p1 = T.p1 != nullptr;
  unsigned long long tests = 0, descent = 0, broot = 0, p1s = 0;
  uint32_t iters = 0, lanes = 0, nsp = 0, handed = 0;
  for (;;) {
    const uint32_t gstart = ctl[0], ng = ctl[1];
    const bool front_done = ctl[2] != 0u;
    // ---- tail: hand the bottom of a deep stack to the next round, 64 items per
    // wave, beyond `keep` items per wave (before this iteration's pops and
    // pushes, so it reads only entries written before the last barrier and
    // never the slots being rewritten)
    if (front_done && (top - bot) > keep * W + NT) {
      const uint32_t s3 = (bot + tid) & (S - 1);
      const Item mine = Item{sa0[s3], sa1[s3], sb0[s3], sb1[s3], sp[s3]};
      const uint32_t sg = (home + W * (nsp++) + w) & (kSegs - 1);
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, 64u);
      b = lane0(b) + lane;
      if (b < segcap) spill[(size_t)sg * segcap + b] = mine;
      else atomicOr(T.overflow, 1u);
      bot += NT;
      handed += NT;
    }
    const uint32_t ns = min(NT, top - bot);
    if (ns + ng == 0u && front_done) break;  // block-uniform: ctl and the stack bounds are equal in every thread
    // ---- batch: the top ns stack items, then ng frontier items
    Item it{0, 0, 0, 0, 0};
    if (tid < ns) {
      const uint32_t s = (top - ns + tid) & (S - 1);
      it = Item{sa0[s], sa1[s], sb0[s], sb1[s], sp[s]};
    } else if (tid < ns + ng) {
      it = in[gstart + tid - ns];
    }
    top -= ns;
    const uint32_t n = ns + ng;
    Expansion x;
    x.push = x.leaf = x.surf = 0;
    x.tests = x.descent = x.broot = 0;
    if (tid < n) expand_item<K>(T.nodes, it, x);
    tests += x.tests; descent += x.descent; broot += x.broot; p1s += __popc(x.leaf);
    ++iters;
    lanes += n;
    uint32_t wsum;
    const uint32_t kpre = wave_prefix_small(__popc(x.push), &wsum);
    if (lane == 0) wtot[w] = wsum;
    // ---- leaf pairs: stage in this wave's LDS (all of P1 when it is kept, else only P2)
    const uint32_t lm = keep_p1 ? x.leaf : x.surf;
    uint32_t ltot;
    const uint32_t lpre = wave_prefix_small(__popc(lm), &ltot);
    if (ltot) {
      uint32_t o = nstage + lpre;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((lm >> k) & 1u) {
          lp_pair[w][o] = it.pair | (((x.surf >> k) & 1u) << 31);
          lp_ta[w][o] = x.ta[k >> 1];
          lp_tb[w][o] = x.tb[k & 1];
          ++o;
        }
      nstage += ltot;
    }
    __syncthreads();  // A: the stack reads are done and every wave's push total is in wtot
    uint32_t woff = 0, total = 0;
#pragma unroll
    for (int v = 0; v < W; ++v) {
      const uint32_t t = wtot[v];
      woff += (uint32_t)v < w ? t : 0u;
      total += t;
    }
    // ---- children: the shared stack when they fit, else the next round (per wave)
    if (total) {
      if ((top - bot) + total <= S) {
        uint32_t o = top + woff + kpre;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((x.push >> k) & 1u) {
            const uint32_t s2 = (o++) & (S - 1);
            const Item& t = x.kids[k];
            sa0[s2] = t.a0; sa1[s2] = t.a1; sb0[s2] = t.b0; sb1[s2] = t.b1; sp[s2] = t.pair;
          }
        top += total;
      } else {
        const uint32_t sg = (home + W * (nsp++) + w) & (kSegs - 1);
        const uint32_t mine = wtot[w];
        if (mine) {
          uint32_t b = 0;
          if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, mine);
          b = lane0(b);
          emit_items(x, spill + (size_t)sg * segcap, b + kpre, segcap, T.overflow);
        }
      }
    }
    if (nstage > SB - 256) {
      wave_lds_sync();
      flush_leaves(T, lp_pair[w], lp_ta[w], lp_tb[w], nstage);
      nstage = 0;
    }
    if (w == 0) {
      ++wave0_iters;
      plan(min(NT, top - bot));
    }
    __syncthreads();  // B: pushes, hand-offs and the next grab are visible to every wave
  }
  wave_lds_sync();
  if (nstage) flush_leaves(T, lp_pair[w], lp_ta[w], lp_tb[w], nstage);
  CounterStripe* st = stripe_of(T.stripes, wave_global());
  const unsigned long long wtests = wave_sum64(tests);
  if (lane == 0 && wtests) {
    atomicAdd(&st->tests, wtests);
    atomicAdd(&st->ptests, wtests);
  }
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
  if (tid == 0) {  // schedule statistics per block: iterations, items expanded, grabs, hand-offs
    atomicAdd(&st->iters, (unsigned long long)iters);
    atomicAdd(&st->lanes, (unsigned long long)lanes);
    atomicAdd(&st->grabs, (unsigned long long)grabs);
    if (handed) atomicAdd(&st->handoffs, (unsigned long long)handed);
    atomicMax(&st->max_iters, (unsigned long long)iters);
  }
}

// The shared-stack round with 4 lanes per work item (expand_quad): a block of
// W waves expands up to 16 W IPQ items per iteration, each quad IPQ items
// (slot q of quad g takes the iteration's item g + 16 W q), loading their
// records in coalesced 64-byte runs.  The stack holds S = 2^SL items.  Wave 0
// claims frontier chunks one iteration ahead and puts the claimed items on
// the stack, so an iteration only pops the stack: the claim's round trip and
// the items' load overlap the expansion instead of preceding it.  The two
// barriers per iteration, the tail hand-off, spills and the per-wave leaf
// staging work as in k_persist_blk.
//
// COUNT (record-counting calls, never timed): the quad's first lane also marks
// every record its item reads in T.touched, so the host can count the
// distinct records a round reads.  budget: iterations after which a block
// whose frontier is gone hands its whole stack on (~0u: never).
//
// SLAB (KDOP_OPT_NODE_SLAB): a wave stages the records of its 16 items in a
// node slab in LDS first -- eight consecutive lanes load one record as whole
// 16-byte words, so a wave-level load covers 128 / rec bytes records in full
// lines -- and each quad then reads its words from the slab.  The slab takes
// 64 records x rec_stride(K) floats per wave (8 KB at K >= 18), which the LDS
// holds for fewer workgroups per CU.
template <int K, int W, int SL, bool COUNT, int IPQ, bool SLAB = false>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(7))) void k_persist_quad(TravArgs T, Item* __restrict__ buf0, Item* __restrict__ buf1, uint32_t cap,
                                                         CallState* __restrict__ C, uint32_t round, uint32_t keep, uint32_t last,
                                                         uint32_t budget) {
  constexpr uint32_t NT = 64u * W;          // threads
  constexpr uint32_t QPI = 16u * W;         // quads
  constexpr uint32_t IPI = QPI * IPQ;       // items per iteration
  constexpr uint32_t S = 1u << SL;
  static_assert(S >= 2u * NT && S >= 2u * IPI, "stack holds a full iteration and a hand-off");
  static_assert(S >= 6u * IPI, "stack holds its remainder, an iteration's children and a grab");
  static_assert(kSegs == 64, "one lane of wave 0 per segment");
  static_assert(IPQ == 1 || IPQ == 2, "one or two items per quad");
  constexpr uint32_t SB = 64u * IPQ + 64u;  // staged leaf pairs per wave; flushed above SB - 64 IPQ (one iteration adds <= 64 IPQ)
  __shared__ uint32_t sa0[S], sa1[S], sb0[S], sb1[S], sp[S];
  __shared__ uint32_t lp_pair[W][SB], lp_ta[W][SB], lp_tb[W][SB];
  __shared__ uint32_t xch[W + 2];  // per iteration: each wave's pushes, then wave 0's grabbed items and frontier-done flag
  static_assert(!SLAB || IPQ == 1, "node slabs with one item per quad");
  constexpr uint32_t RW = rec_stride(K) / 4;  // 16-byte words per record
  __shared__ float4 slab[SLAB ? W : 1][SLAB ? 64u * RW : 1u];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = uniform(tid >> 6);
  const uint32_t gi = w * 16u + (lane >> 2);  // this quad's item of an iteration (slot 0)
  const float4* __restrict__ qbase = quad_lane_base<K>(T.nodes, lane & 3u);
  uint32_t final_level = C->final_level, final_count = 0;
  if (round == 0) {
    const bool recorded = C->final_set != 0u;
    if (!recorded) final_level = last;
    final_count = recorded ? C->final_count : min(C->level_fill[last], cap);
    if (!recorded && blockIdx.x == 0 && tid == 0) {
      C->final_level = final_level;
      C->final_count = final_count;
    }
  }
  const bool odd = ((final_level + round) & 1u) != 0u;
  const Item* __restrict__ in = odd ? buf1 : buf0;
  Item* __restrict__ spill = odd ? buf0 : buf1;
  const uint32_t segcap = cap / kSegs;
  const uint32_t bps = max(1u, gridDim.x / (uint32_t)kSegs);  // blocks per segment
  const uint32_t home = blockIdx.x & (kSegs - 1);
  // wave 0 only: lane s holds segment s of this round's frontier
  uint32_t seg_base = 0, seg_len = 0, cur = home, next_est = 0;
  unsigned long long known = ~0ull;
  bool wfront_done = true;  // wave 0's view
  if (w == 0) {
    if (round == 0) {
      const uint32_t n = min(final_count, cap);
      seg_base = (uint32_t)(((unsigned long long)n * lane) / kSegs);
      seg_len = (uint32_t)(((unsigned long long)n * (lane + 1)) / kSegs) - seg_base;
    } else {
      seg_base = lane * segcap;
      seg_len = min(C->seg[round][lane].fill, segcap);
    }
    known = __ballot(seg_len == 0u);
    wfront_done = known == ~0ull;
    if (!wfront_done && ((known >> cur) & 1ull)) cur = next_open(known, cur);
  }
  uint32_t grabs = 0, wave0_iters = 0;
  // Wave 0 takes the frontier one iteration ahead.  grab_issue claims the next
  // chunk of its current segment before the expansion's loads, so the claim's
  // round trip overlaps them; grab_finish (after the tests) loads the claimed
  // items, and they go on the shared stack on top of the children, where the
  // next iteration takes them like any stack item.  rem: items left on the
  // stack after this iteration's pops; a block claims only while rem < IPI,
  // so the claims are guided by the stack it holds, and at most 64 items (one
  // per lane of wave 0) at a time.
  uint32_t g_want = 0, g_len = 0, g_base = 0, g_raw = 0;
  bool g_on = false;
  auto grab_issue = [&](uint32_t rem) {
    cur = uniform(cur); next_est = uniform(next_est);
    g_on = false;
    if (rem < IPI && !wfront_done) {
      g_len = __builtin_amdgcn_readlane(seg_len, cur);
      g_base = __builtin_amdgcn_readlane(seg_base, cur);
      const uint32_t remaining = g_len > next_est ? g_len - next_est : 0u;
      const uint32_t want = remaining / (2u * bps);
      // at most one item per lane of wave 0, which loads and stacks them
      const uint32_t most = min(IPI - rem, 64u);
      g_want = want < 1u ? 1u : (want > most ? most : want);
      g_raw = 0;
      if (lane == 0) g_raw = atomicAdd(&C->seg[round][cur].next, g_want);
      g_on = true;
    }
  };
  // The claimed items (lane i < ng loads item i into fit), the segment
  // bookkeeping and the frontier state; returns ng.
  auto grab_finish = [&](Item& fit) -> uint32_t {
    uint32_t ng = 0;
    bool refresh = false;
    if (g_on) {
      const uint32_t start = lane0(g_raw);
      if (start < g_len) {
        ++grabs;
        next_est = start + g_want;
        ng = min(g_want, g_len - start);
        if (lane < ng) fit = in[g_base + start + lane];
        if (start + g_want >= g_len) {  // this block took the segment's last chunk
          if (lane == 0) atomicOr(&C->done[round], 1ull << cur);
          known |= 1ull << cur;
        }
      } else {
        known |= 1ull << cur;
        refresh = true;
      }
    } else if (!wfront_done && (wave0_iters & 15u) == 15u) {
      refresh = true;  // a block living off its stack looks at the frontier now and then
    }
    if (refresh) {
      unsigned long long d = 0;
      if (lane == 0) d = __hip_atomic_load(&C->done[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      known |= ((unsigned long long)lane0((uint32_t)(d >> 32)) << 32) | lane0((uint32_t)d);
    }
    if (!wfront_done && ((known >> cur) & 1ull)) {
      if (known == ~0ull) wfront_done = true;
      else { cur = next_open(known, cur); next_est = 0; }
    }
    ++wave0_iters;
    return ng;
  };
  auto put = [&](uint32_t s, const Item& t) {
    sa0[s] = t.a0; sa1[s] = t.a1; sb0[s] = t.b0; sb1[s] = t.b1; sp[s] = t.pair;
  };
  if (w == 0) {  // the first chunk, straight onto the stack
    grab_issue(0u);
    Item fit{0, 0, 0, 0, 0};
    const uint32_t ng = grab_finish(fit);
    if (lane < ng) put(lane, fit);
    if (lane == 0) {
      xch[W] = ng;
      xch[W + 1] = wfront_done ? 1u : 0u;
    }
  }
  __syncthreads();
  uint32_t top = xch[W], bot = 0;  // stack = [bot, top) modulo S, equal in every thread
  bool front_done = xch[W + 1] != 0u;
  uint32_t nstage = 0;        // this wave's staged leaf pairs
  const bool keep_p1 = T.p1 != nullptr;
  uint32_t tests = 0, descent = 0, broot = 0, p1s = 0;
  uint32_t iters = 0, lanes = 0, nsp = 0, handed = 0, recs = 0;
  QuadRec<K> A0[IPQ], A1[IPQ], B0[IPQ], B1[IPQ];  // this lane's words of each slot's records
  for (;;) {
    // ---- tail: hand the bottom of a deep stack to the next round, 64 items per
    // wave, beyond `keep` items per wave (before this iteration's pops and pushes)
    if (front_done && (top - bot) > keep * W + NT) {
      const uint32_t s3 = (bot + tid) & (S - 1);
      const Item mine = Item{sa0[s3], sa1[s3], sb0[s3], sb1[s3], sp[s3]};
      const uint32_t sg = (home + W * (nsp++) + w) & (kSegs - 1);
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, 64u);
      b = lane0(b) + lane;
      if (b < segcap) spill[(size_t)sg * segcap + b] = mine;
      else atomicOr(T.overflow, 1u);
      bot += NT;
      handed += NT;
    }
    // ---- budget: once the frontier is gone, a block that has run `budget`
    // iterations hands its whole stack to the next round and stops, so a deep
    // subtree is spread over every block there instead of running on here
    if (front_done && iters >= budget && top != bot) {
      for (uint32_t c0 = 0; c0 < top - bot; c0 += NT) {
        const uint32_t n = min(NT, top - bot - c0);
        const uint32_t nw = n > 64u * w ? min(64u, n - 64u * w) : 0u;  // this wave's share
        if (nw) {
          const uint32_t sg = (home + W * (nsp++) + w) & (kSegs - 1);
          uint32_t b = 0;
          if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, nw);
          b = lane0(b) + lane;
          if (lane < nw) {
            const uint32_t s3 = (bot + c0 + tid) & (S - 1);
            if (b < segcap) spill[(size_t)sg * segcap + b] = Item{sa0[s3], sa1[s3], sb0[s3], sb1[s3], sp[s3]};
            else atomicOr(T.overflow, 1u);
          }
        }
      }
      handed += top - bot;
      bot = top;
    }
    const uint32_t ns = min(IPI, top - bot);
    if (ns == 0u && front_done) break;  // block-uniform
    if (w == 0) grab_issue(top - bot - ns);
    // ---- batch: the top ns stack items; slot q of this quad
    Item it[IPQ];
    bool live[IPQ];
#pragma unroll
    for (int q = 0; q < IPQ; ++q) {
      const uint32_t g = gi + (uint32_t)q * QPI;
      it[q] = Item{0, 0, 0, 0, 0};
      live[q] = g < ns;
      if (live[q]) {
        const uint32_t s = (top - ns + g) & (S - 1);
        it[q] = Item{sa0[s], sa1[s], sb0[s], sb1[s], sp[s]};
      }
      if (live[q] && (lane & 3u) == 0u) {
        const uint32_t na = (it[q].a0 & I_EXP) ? 2u : 1u, nb = (it[q].b0 & I_EXP) ? 2u : 1u;
        recs += na + nb;
        if (COUNT) {
          T.touched[it[q].a0 & IDX_MASK] = 1u;
          if (na == 2u) T.touched[it[q].a1] = 1u;
          T.touched[it[q].b0 & IDX_MASK] = 1u;
          if (nb == 2u) T.touched[it[q].b1] = 1u;
        }
      }
    }
    top -= ns;
    QuadOut o[IPQ];
    if constexpr (SLAB) {
      // lane 4 q + s names record slot s of quad q's item, so record rho of the
      // wave's slab is named by lane rho
      const uint32_t mine = quad_record(it[0], lane & 3u);
      const float4* __restrict__ nodes4 = reinterpret_cast<const float4*>(T.nodes);
      float4* __restrict__ sl = slab[w];
#pragma unroll
      for (uint32_t k = 0; k < RW; ++k) {
        const uint32_t x = k * 64u + lane, rho = x / RW;
        const uint32_t rec = (uint32_t)__shfl((int)mine, (int)rho, 64);
        sl[x] = nodes4[(size_t)rec * RW + (x % RW)];
      }
      wave_lds_sync();
      const uint32_t r = lane & 3u, qb = (lane & ~3u) * RW;  // this quad's first slab word
#pragma unroll
      for (int t = 0; t < Quad<K>::NQ; ++t) {
        const uint32_t wd = min(r + 4u * (uint32_t)t, RW - 1u);
        A0[0].w[t] = sl[qb + wd];
        A1[0].w[t] = sl[qb + RW + wd];
        B0[0].w[t] = sl[qb + 2u * RW + wd];
        B1[0].w[t] = sl[qb + 3u * RW + wd];
      }
      wave_lds_sync();
      expand_quad_test<K>(it[0], live[0], A0[0], A1[0], B0[0], B1[0], o[0]);
    } else {
#pragma unroll
      for (int q = 0; q < IPQ; ++q) expand_quad<K>(qbase, it[q], live[q], A0[q], A1[q], B0[q], B1[q], o[q]);
    }
    ++iters;
    lanes += ns;
    unsigned long long pm[IPQ];
    uint32_t kpre[IPQ], wsum = 0;
#pragma unroll
    for (int q = 0; q < IPQ; ++q) {  // pushes in slot order, lanes in order within a slot
      tests += o[q].test; descent += o[q].descent; broot += o[q].broot; p1s += o[q].leaf;
      pm[q] = __ballot(o[q].push);
      kpre[q] = wsum + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm[q] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm[q], 0u));
      wsum += (uint32_t)__popcll(pm[q]);
    }
    // ---- leaf pairs: stage in this wave's LDS (all of P1 when it is kept, else only P2)
#pragma unroll
    for (int q = 0; q < IPQ; ++q) {
      const bool lst = keep_p1 ? o[q].leaf : o[q].surf;
      const unsigned long long lm = __ballot(lst);
      if (lm) {
        if (lst) {
          const uint32_t o2 = nstage + __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u));
          lp_pair[w][o2] = o[q].kid.pair | (o[q].surf ? 0x80000000u : 0u);
          lp_ta[w][o2] = o[q].ta;
          lp_tb[w][o2] = o[q].tb;
        }
        nstage += (uint32_t)__popcll(lm);
      }
    }
    // ---- the next frontier chunk (wave 0): its claim has returned by now
    Item fit{0, 0, 0, 0, 0};
    if (w == 0) {
      const uint32_t ng = grab_finish(fit);
      if (lane == 0) {
        xch[W] = ng;
        xch[W + 1] = wfront_done ? 1u : 0u;
      }
    }
    if (lane == 0) xch[w] = wsum;
    __syncthreads();  // A: the stack reads are done; every wave's push total and the grab are in xch
    uint32_t woff = 0, total = 0;
#pragma unroll
    for (int v = 0; v < W; ++v) {
      const uint32_t t = xch[v];
      woff += (uint32_t)v < w ? t : 0u;
      total += t;
    }
    const uint32_t ng = xch[W];
    front_done = xch[W + 1] != 0u;
    // ---- children: the shared stack when they fit, else the next round (per wave)
    if (total) {
      if ((top - bot) + total <= S) {
#pragma unroll
        for (int q = 0; q < IPQ; ++q)
          if (o[q].push) put((top + woff + kpre[q]) & (S - 1), o[q].kid);
        top += total;
      } else {
        const uint32_t sg = (home + W * (nsp++) + w) & (kSegs - 1);
        if (wsum) {
          uint32_t b = 0;
          if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, wsum);
          b = lane0(b);
#pragma unroll
          for (int q = 0; q < IPQ; ++q)
            if (o[q].push) {
              if (b + kpre[q] < segcap) spill[(size_t)sg * segcap + b + kpre[q]] = o[q].kid;
              else atomicOr(T.overflow, 1u);
            }
        }
      }
    }
    // ---- the grabbed frontier items on top of the children (they always fit:
    // a block claims only while fewer than IPI items stay on its stack, and
    // S >= 6 IPI; the spill is a guard)
    if (ng) {
      if ((top - bot) + ng <= S) {
        if (w == 0 && lane < ng) put((top + lane) & (S - 1), fit);
        top += ng;
      } else if (w == 0) {
        const uint32_t sg = (home + W * (nsp++)) & (kSegs - 1);
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&C->seg[round + 1][sg].fill, ng);
        b = lane0(b) + lane;
        if (lane < ng) {
          if (b < segcap) spill[(size_t)sg * segcap + b] = fit;
          else atomicOr(T.overflow, 1u);
        }
      }
    }
    if (nstage > SB - 64u * IPQ) {
      wave_lds_sync();
      flush_leaves(T, lp_pair[w], lp_ta[w], lp_tb[w], nstage);
      nstage = 0;
    }
    __syncthreads();  // B: pushes, hand-offs and the grabbed items are visible to every wave
  }
  wave_lds_sync();
  if (nstage) flush_leaves(T, lp_pair[w], lp_ta[w], lp_tb[w], nstage);
  CounterStripe* st = stripe_of(T.stripes, wave_global());
  const unsigned long long wtests = wave_sum64(tests);
  if (lane == 0 && wtests) {
    atomicAdd(&st->tests, wtests);
    atomicAdd(&st->ptests, wtests);
  }
  wave_add64(&st->descent, descent);
  wave_add64(&st->branch_overlaps, broot);
  wave_add64(&st->p1, p1s);
  wave_add64(&st->records, recs);
  if (tid == 0) {  // schedule statistics per block: iterations, items expanded, grabs, hand-offs
    atomicAdd(&st->iters, (unsigned long long)iters);
    atomicAdd(&st->lanes, (unsigned long long)lanes);
    atomicAdd(&st->grabs, (unsigned long long)grabs);
    if (handed) atomicAdd(&st->handoffs, (unsigned long long)handed);
    atomicMax(&st->max_iters, (unsigned long long)iters);
  }
}

// The spill left in round `round`'s segments becomes breadth-first level
// `level` (one 64-lane block, before the level launches): the persistent
// rounds spill into kSegs segments of cap / kSegs slots, the layout of a
// segmented level, so the fills are copied over and the levels strategy
// finishes the traversal from there.  level >= 1 (segmented input) and of the
// parity that selects the buffer the spill is in; the level's output set and
// the next are cleared, and the stop flag is reset for the levels to run.
__global__ __launch_bounds__(64) void k_spill_to_level(CallState* __restrict__ C, uint32_t round, uint32_t level) {
  const uint32_t lane = threadIdx.x;
  C->lseg[level % 3u][lane].fill = C->seg[round][lane].fill;
  C->lseg[(level + 1u) % 3u][lane].fill = 0u;
  C->lseg[(level + 2u) % 3u][lane].fill = 0u;
  if (lane == 0u) C->final_set = 0u;
}

// Fold the striped counters and the segment fills into the call summary, one
// 1024-thread block: wave f < kStripeFields reduces counter field f over the
// stripes (lane s reads stripe s), waves 14 and 15 the segment fills of four
// rounds each, so every load is independent and each reduction is one
// wave's.  The results meet in LDS, and wave 0 alone writes the summary
// straight to pinned host memory (`host`, mapped, coherent), makes its stores
// visible with one system-scope fence, and writes the `ready` word last, so
// the host waits for that word instead of a copy and a stream
// synchronisation.  (A system-scope fence writes the L2 back; one wave
// fencing instead of all sixteen takes the fold from 9.0 to 6.0 us.)  With
// final_round >= -1 and a complete call (no overflow, nothing spilled into
// final_round's segments; -1: levels strategy, frontier found empty) the
// block then zeroes the call state, so the next call needs no memset.
constexpr int kSumThreads = 1024;
__global__ __launch_bounds__(kSumThreads) void k_sum_counters(CallState* __restrict__ C, uint32_t cap, int final_round,
                                                              CallSummary* __restrict__ host) {
  static_assert(kStripes == 64 && kSegs == 64, "one lane per stripe and per segment");
  static_assert(kStripeFields == 14 && kMaxRounds == 8, "14 field waves, 2 round waves");
  static_assert(sizeof(CounterStripe) * kStripes == 16 * kSumThreads, "one 16-byte clear per thread");
  __shared__ unsigned long long total_s[kStripeFields];
  __shared__ uint32_t round_total_s[kMaxRounds], round_maxseg_s[kMaxRounds];
  __shared__ uint32_t pending_s;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  CallSummary& S = *host;
  if (tid == 0) pending_s = 0u;
  __syncthreads();
  if (w < (uint32_t)kStripeFields) {
    unsigned long long x = reinterpret_cast<const unsigned long long*>(&C->stripe[lane])[w];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const unsigned long long y = __shfl_xor(x, d, 64);
      x = (w == (uint32_t)kStripeFields - 1u) ? max(x, y) : x + y;
    }
    if (lane == 0) total_s[w] = x;
  } else {
    const uint32_t segcap = cap / kSegs;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int r = 4 * (int)(w - 14u) + h;
      const uint32_t f = C->seg[r][lane].fill;
      uint32_t tot = min(f, segcap), mx = f;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        tot += __shfl_xor(tot, d, 64);
        mx = max(mx, (uint32_t)__shfl_xor(mx, d, 64));
      }
      if (lane == 0) {
        round_total_s[r] = r == 0 ? C->final_count : tot;
        round_maxseg_s[r] = mx;
        if (r == final_round) pending_s = tot;
      }
    }
  }
  __syncthreads();
  const uint32_t ovf = C->overflow;
  const bool reset = final_round >= -1 && ovf == 0u && pending_s == 0u && (final_round >= 0 || C->final_set != 0u);
  if (w == 0) {
    if (lane < (uint32_t)kStripeFields) S.total[lane] = total_s[lane];
    if (lane < (uint32_t)kMaxRounds) {
      S.round_total[lane] = round_total_s[lane];
      S.round_maxseg[lane] = round_maxseg_s[lane];
    }
    S.level_fill[lane] = C->level_fill[lane];
    if (lane == 0) {
      S.overflow = ovf;
      S.fills = C->fills;
      S.final_level = C->final_level;
      S.final_count = C->final_count;
      S.final_set = C->final_set;
      S.reset = reset ? 1u : 0u;
    }
    __threadfence_system();  // the wave's summary stores are visible to the host
    if (lane == 0) __hip_atomic_store(&S.ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();  // every read of the state is done before it is cleared
  if (!reset) return;
  reinterpret_cast<float4*>(C->stripe)[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (tid < (uint32_t)(kMaxRounds * kSegs)) {
    SegCtr& sc = C->seg[tid / kSegs][tid % kSegs];
    sc.fill = 0u;
    sc.next = 0u;
  }
  if (tid < 64u) C->level_fill[tid] = 0u;
  if (tid < 3u * (uint32_t)kSegs) C->lseg[tid / kSegs][tid % kSegs].fill = 0u;
  if (tid < (uint32_t)kMaxRounds) C->done[tid] = 0ull;
  if (tid == 0) {
    C->overflow = 0u;
    C->fills = Fills{0u, 0u, 0u, 0u, 0u};
    C->final_level = 0u;
    C->final_count = 0u;
    C->final_set = 0u;
  }
}

}  // namespace kdop_amd
