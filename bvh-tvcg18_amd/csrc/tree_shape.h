// tree_shape.h -- the reference's chunked-tree shape, host side.
//
// kdop::make_tree splits the (Morton ordered) tetrahedra into chunks of L
// leaves, where L comes from how many kdop::Node<float,K> fit in mem_bytes/2
// (SIMULATION/KDOP/KDOP/include/kdop_make_tree.h:216-229), then stacks H-1
// levels of super chunks of L chunks each (:255-283).  The arithmetic mixes
// size_t, float (VT::half, VT::one) and double (log); it is kept as written so
// the shapes (and so the reference's node-pair test counts) come out equal.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace kdop_amd {

struct Shape {
  uint64_t node_bytes, n_perfect, L, C, H;
};

// tiny_power2.h:35-62
inline uint64_t upper_pow2(uint64_t v) {
  uint64_t p = 1;
  for (int i = 0; i < 32; ++i) {
    if (p >= v) return p;
    p <<= 1;
  }
  return 0;
}
inline uint64_t lower_pow2(uint64_t v) {
  uint64_t p = uint64_t(1) << 30;
  for (int i = 0; i < 32; ++i) {
    if (p <= v) return p;
    p >>= 1;
  }
  return 0;
}

// Returns false when no two-leaf chunk fits (the reference would underflow).
inline bool chunk_shape(int K, uint64_t mem_bytes, uint64_t M, Shape& s) {
  if (M == 0) return false;
  s.node_bytes = (uint64_t)(4 * K) + 3 * 8;  // DOP<float,K> + parent/start/end size_t (LP64)
  const uint64_t n_max = (uint64_t)std::floor((float)(mem_bytes / s.node_bytes) * 0.5f);
  const uint64_t m = std::min(lower_pow2(n_max), upper_pow2(M * 2));
  if (m < 2) return false;
  s.n_perfect = m - 1;
  s.L = (uint64_t)((float)(s.n_perfect + 1) * 0.5f);
  if (s.L < 1) return false;
  s.C = (uint64_t)std::ceil(1.0f * (float)M / (float)s.L);
  // log(1)/log(L) is 0 for L > 1, and the one-leaf case L == 1 only arises
  // with C == 1: both give a single level.
  s.H = (s.C <= 1) ? 1 : (uint64_t)std::max(std::ceil(std::log((double)s.C) / std::log((double)s.L)), 1.0);
  return true;
}

// Chunks above level h+1: ceil(1.0f * C / L) in float, as :258 computes it.
inline uint64_t chunks_above(uint64_t C, uint64_t L) { return (uint64_t)std::ceil(1.0f * (float)C / (float)L); }

}  // namespace kdop_amd
