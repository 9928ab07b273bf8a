// meshgen.cpp -- host-side mesh preparation (box-grid tetrahedraliser, Morton
// reorder, surface map, chunked-tree shape, seeded random stream).  Pure C++,
// no device code; compiled into libkdop_amd.so.
#include "kdop_meshgen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace {

// ------------------------------------------------------------ random stream
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro {
  uint64_t s[4];
  Xoshiro(uint64_t seed, uint64_t stream) {
    uint64_t sm = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(sm);
  }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {  // Box-Muller, one value per call (the sine partner is dropped)
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// ------------------------------------------------------------- box grid
// Corner c of a cube has offsets (c&1, (c>>1)&1, (c>>2)&1).  Even cubes keep
// the central tetrahedron {1,2,4,7} and cut corners 0,3,5,6; odd cubes keep
// {0,3,5,6} and cut corners 1,2,4,7, so shared faces get the same diagonal.
const int EVEN_TETS[5][4] = {{0, 1, 2, 4}, {3, 2, 1, 7}, {5, 4, 7, 1}, {6, 7, 4, 2}, {1, 2, 4, 7}};
const int ODD_TETS[5][4] = {{1, 0, 5, 3}, {2, 3, 6, 0}, {4, 5, 0, 6}, {7, 6, 3, 5}, {0, 3, 5, 6}};

inline double signed_vol6(const double* a, const double* b, const double* c, const double* d) {
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return w[0] * (u[1] * v[2] - u[2] * v[1]) + w[1] * (u[2] * v[0] - u[0] * v[2]) + w[2] * (u[0] * v[1] - u[1] * v[0]);
}

}  // namespace

extern "C" {

void kdop_box_mesh_size(uint32_t n, uint32_t* nv, uint32_t* nt) {
  *nv = (n + 1) * (n + 1) * (n + 1);
  *nt = 5 * n * n * n;
}

int kdop_make_box_mesh(uint32_t n, float size, double jitter, uint64_t seed, float* X, float* Y, float* Z, uint32_t* tets) {
  if (n == 0 || !(size > 0.0f)) return -1;
  const uint32_t m = n + 1;
  const size_t nv = (size_t)m * m * m;
  const double h = (double)size / n;
  std::vector<double> P(3 * nv);
  auto vid = [m](uint32_t i, uint32_t j, uint32_t k) { return (uint32_t)(i + m * (j + m * k)); };
  for (uint32_t k = 0; k < m; ++k)
    for (uint32_t j = 0; j < m; ++j)
      for (uint32_t i = 0; i < m; ++i) {
        const uint32_t v = vid(i, j, k);
        P[3 * v] = -0.5 * size + i * h;
        P[3 * v + 1] = -0.5 * size + j * h;
        P[3 * v + 2] = -0.5 * size + k * h;
      }
  // tets, positively oriented on the undeformed grid
  size_t t = 0;
  for (uint32_t k = 0; k < n; ++k)
    for (uint32_t j = 0; j < n; ++j)
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[8];
        for (int q = 0; q < 8; ++q) c[q] = vid(i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
        const int(*T)[4] = ((i + j + k) & 1) ? ODD_TETS : EVEN_TETS;
        for (int e = 0; e < 5; ++e, ++t) {
          uint32_t a = c[T[e][0]], b = c[T[e][1]], cc = c[T[e][2]], d = c[T[e][3]];
          if (signed_vol6(&P[3 * a], &P[3 * b], &P[3 * cc], &P[3 * d]) < 0.0) std::swap(cc, d);
          tets[4 * t] = a; tets[4 * t + 1] = b; tets[4 * t + 2] = cc; tets[4 * t + 3] = d;
        }
      }
  const size_t nt = t;
  if (jitter > 0.0) {
    // vertex -> incident tets
    std::vector<uint32_t> cnt(nv + 1, 0), inc(4 * nt);
    for (size_t q = 0; q < 4 * nt; ++q) cnt[tets[q] + 1]++;
    for (size_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (size_t q = 0; q < 4 * nt; ++q) inc[fill[tets[q]]++] = (uint32_t)(q / 4);
    std::vector<double> vol0(nt);
    for (size_t e = 0; e < nt; ++e)
      vol0[e] = signed_vol6(&P[3 * tets[4 * e]], &P[3 * tets[4 * e + 1]], &P[3 * tets[4 * e + 2]], &P[3 * tets[4 * e + 3]]);
    Xoshiro rng(seed, 0x6a697474ull);
    const double sigma = jitter * h;
    for (uint32_t k = 1; k < n; ++k)
      for (uint32_t j = 1; j < n; ++j)
        for (uint32_t i = 1; i < n; ++i) {
          const uint32_t v = vid(i, j, k);
          const double o[3] = {P[3 * v], P[3 * v + 1], P[3 * v + 2]};
          bool ok = false;
          for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
            const double dx = rng.normal() * sigma, dy = rng.normal() * sigma, dz = rng.normal() * sigma;
            P[3 * v] = o[0] + dx; P[3 * v + 1] = o[1] + dy; P[3 * v + 2] = o[2] + dz;
            ok = true;
            for (uint32_t r = cnt[v]; r < cnt[v + 1] && ok; ++r) {
              const uint32_t e = inc[r];
              const double vv = signed_vol6(&P[3 * tets[4 * e]], &P[3 * tets[4 * e + 1]], &P[3 * tets[4 * e + 2]], &P[3 * tets[4 * e + 3]]);
              if (!(vv >= 0.1 * vol0[e])) ok = false;
            }
          }
          if (!ok) { P[3 * v] = o[0]; P[3 * v + 1] = o[1]; P[3 * v + 2] = o[2]; }
        }
  }
  for (size_t v = 0; v < nv; ++v) { X[v] = (float)P[3 * v]; Y[v] = (float)P[3 * v + 1]; Z[v] = (float)P[3 * v + 2]; }
  return 0;
}

// kdop_mesh_reorder.h:172-353, with a stable tie order.
int kdop_mesh_reorder(uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt, const uint32_t* tets,
                      float* Xo, float* Yo, float* Zo, uint32_t* tets_o) {
  if (nt == 0 || nv == 0) return -1;
  std::vector<float> cx(nt), cy(nt), cz(nt);
  // centroid = 0.25 (double) * float sum, stored back as float (:193-221)
  float min_x = std::numeric_limits<float>::max(), min_y = min_x, min_z = min_x;
  float max_x = std::numeric_limits<float>::min(), max_y = max_x, max_z = max_x;  // reference initialises max with min() (:197-199)
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t* v = tets + 4 * t;
    cx[t] = (float)(0.25 * (double)(X[v[0]] + X[v[1]] + X[v[2]] + X[v[3]]));
    cy[t] = (float)(0.25 * (double)(Y[v[0]] + Y[v[1]] + Y[v[2]] + Y[v[3]]));
    cz[t] = (float)(0.25 * (double)(Z[v[0]] + Z[v[1]] + Z[v[2]] + Z[v[3]]));
    min_x = std::min(min_x, cx[t]); min_y = std::min(min_y, cy[t]); min_z = std::min(min_z, cz[t]);
    max_x = std::max(max_x, cx[t]); max_y = std::max(max_y, cy[t]); max_z = std::max(max_z, cz[t]);
  }
  // 21 bits per axis, one common integer scale (:226-240)
  const int KB = (int)(sizeof(unsigned long long) * 8) / 3;
  const float top = (float)((1 << KB) - 1);
  auto scale_of = [&](float lo, float hi) -> int {
    const float s = top / (hi - lo);
    if (!(s < 2147483647.0f)) return 2147483647;  // degenerate axis (all centroids equal)
    return (int)s;
  };
  int scale = scale_of(min_x, max_x);
  scale = std::min(scale, scale_of(min_y, max_y));
  scale = std::min(scale, scale_of(min_z, max_z));
  std::vector<unsigned long long> code(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    const unsigned long long x = (unsigned long long)(float)(int)((cx[t] - min_x) * (float)scale);
    const unsigned long long y = (unsigned long long)(float)(int)((cy[t] - min_y) * (float)scale);
    const unsigned long long z = (unsigned long long)(float)(int)((cz[t] - min_z) * (float)scale);
    unsigned long long c = 0;
    for (int k = 0; k < KB; ++k) {
      c |= ((x >> k) & 1ull) << (3 * k);
      c |= ((y >> k) & 1ull) << (3 * k + 1);
      c |= ((z >> k) & 1ull) << (3 * k + 2);
    }
    code[t] = c;
  }
  std::vector<uint32_t> order(nt);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  // first-touch vertex numbering (:295-350)
  std::vector<uint32_t> lookup(nv, 0xFFFFFFFFu);
  uint32_t next = 0;
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t* v = tets + 4 * order[i];
    for (int q = 0; q < 4; ++q) {
      uint32_t& o = lookup[v[q]];
      if (o == 0xFFFFFFFFu) {
        o = next++;
        Xo[o] = X[v[q]]; Yo[o] = Y[v[q]]; Zo[o] = Z[v[q]];
      }
      tets_o[4 * i + q] = o;
    }
  }
  // vertices not used by any tetrahedron are appended in input order
  for (uint32_t v = 0; v < nv; ++v)
    if (lookup[v] == 0xFFFFFFFFu) { lookup[v] = next; Xo[next] = X[v]; Yo[next] = Y[v]; Zo[next] = Z[v]; ++next; }
  return 0;
}

// mesh_array_compute_surface_map.h:41-76: a face is on the surface when no other
// tetrahedron shares its three vertices.
int kdop_compute_surface_map(uint32_t nv, uint32_t nt, const uint32_t* tets, uint8_t* surface) {
  if (nv >= (1u << 21)) return -1;
  static const int OPP[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  std::vector<std::pair<uint64_t, uint32_t>> faces((size_t)4 * nt);
  for (uint32_t t = 0; t < nt; ++t)
    for (int f = 0; f < 4; ++f) {
      uint32_t a = tets[4 * t + OPP[f][0]], b = tets[4 * t + OPP[f][1]], c = tets[4 * t + OPP[f][2]];
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
      faces[(size_t)4 * t + f] = {((uint64_t)a << 42) | ((uint64_t)b << 21) | c, 4 * t + f};
    }
  std::sort(faces.begin(), faces.end());
  std::memset(surface, 0, nt);
  for (size_t i = 0; i < faces.size();) {
    size_t j = i + 1;
    while (j < faces.size() && faces[j].first == faces[i].first) ++j;
    if (j - i == 1) {
      const uint32_t tf = faces[i].second;
      surface[tf / 4] |= (uint8_t)(1u << (tf % 4));
    }
    i = j;
  }
  return 0;
}

void kdop_rng_uniform(uint64_t seed, uint64_t stream, uint64_t n, double* out) {
  Xoshiro r(seed, stream);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.uniform();
}
void kdop_rng_normal(uint64_t seed, uint64_t stream, uint64_t n, double* out) {
  Xoshiro r(seed, stream);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.normal();
}
void kdop_rng_quaternions(uint64_t seed, uint64_t stream, uint64_t n, float* out) {
  Xoshiro r(seed, stream);
  for (uint64_t i = 0; i < n; ++i) {
    const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    out[4 * i + 0] = (float)(a * std::sin(2.0 * M_PI * u2));
    out[4 * i + 1] = (float)(a * std::cos(2.0 * M_PI * u2));
    out[4 * i + 2] = (float)(b * std::sin(2.0 * M_PI * u3));
    out[4 * i + 3] = (float)(b * std::cos(2.0 * M_PI * u3));
  }
}

}  // extern "C"
