// kdop_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's sequential kDOP path (diku-dk/bvh-tvcg18,
// backend tag `kdop::sequential`).  It is the checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// The product library never links, loads or calls anything in this file.
//
// Every function cites the reference file:line it restates.  Build flags
// (oracle/Makefile): g++ -O2 -std=c++17 -ffp-contract=off, no fast-math, so
// every float op rounds exactly as the reference's g++ -std=c++11 build does
// (x86-64 SSE, no FMA contraction).
//
// Parity pins (how each function is checked independently of this file, see
// tests/test_oracle_pins.py):
//   dir_table          closed form 1/sqrt(2), 1/sqrt(3) bits (numpy float32)
//   make_dop (leaf)    K=6: exact AABB of the 4 vertices; K=8..26: numpy float32
//                      projections with the same left-to-right sum
//   refit / unions     every parent contains its children exactly; root DOP =
//                      brute-force min/max over all vertex projections
//   make_tree shape    closed-form chunk counts L/C/H quoted for the workloads
//   tandem traversal   pair set == brute-force all-pairs leaf overlap; test count
//                      == closed form on meshes whose leaves all overlap
//                      (4^d pairs per depth), == 1 for separated bodies,
//                      == 1 + C_A*C_B when only the roots overlap
//   rotate             rotation-matrix product within 4 ulp (numpy float64)
//   SAT                LP feasibility (scipy) on random tetrahedron pairs
//   contacts           analytic depth on axis-aligned configurations; FP64
//                      restatement agrees to 1e-5 relative
//   prox_reduce        hand-worked known answers (key order, the precision
//                      boundary, std::unique's last-kept rule); numpy stable
//                      lexsort + first of each run on grid-valued contacts
//                      (tests/test_reduce_oracle_cpu.py)
// The reference itself is unbuildable in this container (its TINY headers need
// Boost, which is not installed), so no output of the compiled reference pins it.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- small math
// tiny_vector_ops_policy.h:33-45 (inner_prod: ((0 + a0*b0) + a1*b1) + a2*b2),
// :47-54 (cross), :57-75 (norm, unit: v * (1/sqrt(v.v)) when length > 0).
template <class T> struct V3 { T x, y, z; };

template <class T> inline V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) {
  T d = T(0);
  d = d + a.x * b.x;
  d = d + a.y * b.y;
  d = d + a.z * b.z;
  return d;
}
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline V3<T> sub(const V3<T>& a, const V3<T>& b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> add(const V3<T>& a, const V3<T>& b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> mul(const V3<T>& a, T s) { return V3<T>{a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> neg(const V3<T>& a) { return V3<T>{-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> unit(const V3<T>& v) {
  T l = std::sqrt(dot(v, v));
  if (l > T(0)) {
    T inv = T(1) / l;
    return mul(v, inv);
  }
  return v;
}
// std::min / std::max as the reference calls them: min(a,b) = (b < a) ? b : a.
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }
// tiny_vector_functions.h:235-239: norm_1(v) = max(|v_i|) (an infinity norm).
template <class T> inline T norm_1(const V3<T>& v) {
  return smax(smax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
}
// tiny_precision.h:21-24: working_precision = epsilon * 10.
template <class T> inline T working_precision() { return std::numeric_limits<T>::epsilon() * T(10); }

// tiny_quaternion_functions.h:137-152 with tiny_quaternion_ops_policy.h:29-38.
// rotate(q, r) = Im( prod( prod(q, r), conj(q) ) ); q stored (x, y, z, w=real).
template <class T> inline V3<T> rotate(const T q[4], const V3<T>& r) {
  const V3<T> u = mk(q[0], q[1], q[2]);
  const T s = q[3];
  // prod(q, r) = ( -u.r , u x r + r*s )
  const T a = -dot(u, r);
  const V3<T> A = add(cross(u, r), mul(r, s));
  // mul_assign(lhs=(a,A), rhs=(s,-u)): imag = a*B + A*b + A x B
  const V3<T> B = neg(u);
  const T b = s;
  V3<T> out;
  out.x = (a * B.x + A.x * b) + (A.y * B.z - A.z * B.y);
  out.y = (a * B.y + A.y * b) + (A.z * B.x - A.x * B.z);
  out.z = (a * B.z + A.z * b) + (A.x * B.y - A.y * B.x);
  return out;
}

// ---------------------------------------------------------- direction tables
// geometry_direction_table.h:53-121 (make3/4/6 and unions), :123-172 (helper:
// N = K/2 in {3,4,6,7,9,10,13}).
template <class T> inline int dir_table(int K, V3<T>* D) {
  const T o = T(1), z = T(0);
  V3<T> d3[3] = {mk(o, z, z), mk(z, o, z), mk(z, z, o)};
  V3<T> d4[4] = {unit(mk(o, o, o)), unit(mk(o, o, -o)), unit(mk(o, -o, o)), unit(mk(o, -o, -o))};
  V3<T> d6[6] = {unit(mk(o, o, z)), unit(mk(o, -o, z)), unit(mk(o, z, o)),
                 unit(mk(z, o, o)), unit(mk(-o, z, o)), unit(mk(z, -o, o))};
  int n = 0;
  auto put = [&](V3<T>* src, int cnt) { for (int i = 0; i < cnt; ++i) D[n++] = src[i]; };
  switch (K / 2) {
    case 3: put(d3, 3); break;
    case 4: put(d4, 4); break;
    case 6: put(d6, 6); break;
    case 7: put(d3, 3); put(d4, 4); break;
    case 9: put(d3, 3); put(d6, 6); break;
    case 10: put(d4, 4); put(d6, 6); break;
    case 13: put(d3, 3); put(d4, 4); put(d6, 6); break;
    default: return -1;
  }
  return n;
}

// ------------------------------------------------------------------ kDOP tree
static const int KMAX = 26;
struct Dop {
  float lo[KMAX / 2];
  float hi[KMAX / 2];
};
inline void dop_empty(Dop& d, int N) {  // geometry_interval.h:28-31 (lower=+max, upper=-max)
  for (int k = 0; k < N; ++k) { d.lo[k] = std::numeric_limits<float>::max(); d.hi[k] = -std::numeric_limits<float>::max(); }
}
// geometry_make_union.h:9-25
inline void dop_union(Dop& out, const Dop& A, const Dop& B, int N) {
  for (int k = 0; k < N; ++k) { out.lo[k] = smin(A.lo[k], B.lo[k]); out.hi[k] = smax(A.hi[k], B.hi[k]); }
}
// geometry_overlap_dop_dop.h:12-24 + geometry_overlap_interval_interval.h:11-21 (strict)
inline bool dop_overlap(const Dop& A, const Dop& B, int N) {
  for (int k = 0; k < N; ++k) {
    if (A.hi[k] < B.lo[k]) return false;
    if (B.hi[k] < A.lo[k]) return false;
  }
  return true;
}
// geometry_dop.h:48-74 (make_dop over the 4 tet vertices, slab-major loop)
inline void make_dop(Dop& d, const V3<float>* pts, int npts, const V3<float>* D, int N) {
  dop_empty(d, N);
  for (int k = 0; k < N; ++k)
    for (int p = 0; p < npts; ++p) {
      const float proj = dot(D[k], pts[p]);
      d.lo[k] = smin(d.lo[k], proj);
      d.hi[k] = smax(d.hi[k], proj);
    }
}

static const size_t UNDEF = 0xFFFF;  // kdop_tree.h:12
struct Node {                         // kdop_tree.h:14-75
  Dop vol;
  size_t parent = UNDEF, start = UNDEF, end = UNDEF;
  bool is_leaf() const { return start == end; }
  bool is_undefined() const { return parent == UNDEF && start == UNDEF && end == UNDEF; }
};
struct SubTree { std::vector<Node> nodes; size_t height = 0; };  // kdop_tree.h:77-113
struct Tree {                                                       // kdop_tree.h:115-200
  int K = 8, N = 4;
  size_t node_bytes = 0, N_perfect = 0, L = 0, H = 0;
  Dop root;
  std::vector<std::vector<SubTree>> levels;  // levels.back() = branches
  std::vector<V3<float>> D;
};

// tiny_power2.h:35-62
inline size_t upper_power2(size_t val) {
  size_t next = 1u;
  for (unsigned i = 0; i < 32u; ++i) { if (next >= val) return next; next <<= 1u; }
  return 0u;
}
inline size_t lower_power2(size_t val) {
  size_t next = size_t(1) << 30u;
  for (unsigned i = 0; i < 32u; ++i) { if (next <= val) return next; next >>= 1u; }
  return 0u;
}

// kdop_make_tree.h:57-121 / :123-186 (median split; children allocated as a
// pair at free_idx, free_idx+1; left subtree completed before the right one).
inline void make_subtree(size_t parent_idx, size_t& free_idx, size_t first, size_t last, SubTree& st, size_t& height) {
  const size_t Nel = last - first + 1u;
  if (Nel > 1) {
    const size_t left = free_idx, right = free_idx + 1;
    free_idx += 2u;
    st.nodes[parent_idx].start = left;
    st.nodes[parent_idx].end = right;
    const size_t mid = (size_t)std::floor((first + last) / 2.0);
    st.nodes[left].parent = parent_idx;
    st.nodes[right].parent = parent_idx;
    size_t hl = 0, hr = 0;
    make_subtree(left, free_idx, first, mid, st, hl);
    make_subtree(right, free_idx, mid + 1, last, st, hr);
    height = std::max(hl, hr) + 1;
  } else {
    st.nodes[parent_idx].start = first;
    st.nodes[parent_idx].end = first;
    height = 0;
  }
}

// Chunking formula, kdop_make_tree.h:216-229.  Returns false when the
// chunk budget cannot hold a two-leaf chunk (the reference would underflow).
inline bool tree_shape(int K, size_t mem_bytes, size_t M, size_t& node_bytes, size_t& N_perfect, size_t& L, size_t& C, size_t& H) {
  node_bytes = (size_t)(4 * K) + 3 * sizeof(size_t);  // sizeof(kdop::Node<float,K>) on LP64
  const size_t N_max = (size_t)std::floor((float)(mem_bytes / node_bytes) * 0.5f);
  const size_t lp = lower_power2(N_max), up = upper_power2(M * 2);
  const size_t mn = std::min(lp, up);
  if (mn < 2 || M == 0) return false;
  N_perfect = mn - 1;
  L = (size_t)((float)(N_perfect + 1) * 0.5f);
  if (L < 1) return false;
  C = (size_t)std::ceil(1.0f * (float)M / (float)L);
  if (C <= 1) {
    H = 1;  // log(1)/log(L) = 0 (L>1) -> max(ceil(0),1) = 1; also covers L==1
  } else {
    H = (size_t)std::max(std::ceil(std::log((double)C) / std::log((double)L)), 1.0);
  }
  return true;
}

// kdop_refit_tree.h:23-80 (branch chunk: reverse sweep, leaves make_dop, internal union)
inline void refit_branch(SubTree& st, const Tree& t, const uint32_t* tets, const float* X, const float* Y, const float* Z) {
  const size_t Nn = st.nodes.size();
  for (size_t i = 0; i < Nn; ++i) {
    Node& nd = st.nodes[Nn - i - 1u];
    if (nd.is_undefined()) continue;
    if (nd.is_leaf()) {
      const uint32_t* tv = tets + 4 * nd.start;
      V3<float> pts[4];
      for (int v = 0; v < 4; ++v) pts[v] = mk(X[tv[v]], Y[tv[v]], Z[tv[v]]);
      make_dop(nd.vol, pts, 4, t.D.data(), t.N);
    } else {
      dop_union(nd.vol, st.nodes[nd.start].vol, st.nodes[nd.end].vol, t.N);
    }
  }
}
// kdop_refit_tree.h:82-115 (super chunk: leaves copy the child chunk's root)
inline void refit_super(SubTree& st, const std::vector<SubTree>& children, int N) {
  const size_t Nn = st.nodes.size();
  for (size_t i = 0; i < Nn; ++i) {
    Node& nd = st.nodes[Nn - i - 1u];
    if (nd.is_undefined()) continue;
    if (nd.is_leaf()) nd.vol = children[nd.start].nodes[0].vol;
    else dop_union(nd.vol, st.nodes[nd.start].vol, st.nodes[nd.end].vol, N);
  }
}
// kdop_refit_tree.h:119-162
inline void refit_tree(Tree& t, const uint32_t* tets, const float* X, const float* Y, const float* Z) {
  for (auto& st : t.levels.back()) refit_branch(st, t, tets, X, Y, Z);
  for (size_t h = t.levels.size() - 1; h >= 1; --h)
    for (auto& st : t.levels[h - 1]) refit_super(st, t.levels[h], t.N);
  t.root = t.levels[0][0].nodes[0].vol;
  for (size_t c = 1; c < t.levels[0].size(); ++c) dop_union(t.root, t.root, t.levels[0][c].nodes[0].vol, t.N);
}

// kdop_make_tree.h:190-288
inline Tree* make_tree(int K, size_t mem_bytes, size_t M, const uint32_t* tets, const float* X, const float* Y, const float* Z) {
  Tree* t = new Tree();
  t->K = K;
  t->N = K / 2;
  t->D.resize(13);
  if (dir_table<float>(K, t->D.data()) < 0) { delete t; return nullptr; }
  size_t C = 0;
  if (!tree_shape(K, mem_bytes, M, t->node_bytes, t->N_perfect, t->L, C, t->H)) { delete t; return nullptr; }
  const size_t L = t->L;
  t->levels.resize(t->H);
  t->levels.back().resize(C);
  for (size_t c = 0; c < C; ++c) {
    SubTree& st = t->levels.back()[c];
    const size_t first = c * L, last = std::min(first + L - 1u, M - 1u);
    st.nodes.resize(t->N_perfect);
    size_t free_idx = 1u;
    make_subtree(0u, free_idx, first, last, st, st.height);
  }
  for (size_t h = t->H - 1; h >= 1; --h) {
    C = (size_t)std::ceil(1.0f * (float)C / (float)L);
    t->levels[h - 1].resize(C);
    for (size_t c = 0; c < C; ++c) {
      SubTree& st = t->levels[h - 1][c];
      const size_t first = c * L, last = std::min(first + L - 1, t->levels[h].size() - 1);
      st.nodes.resize(t->N_perfect);
      size_t free_idx = 1u;
      make_subtree(0u, free_idx, first, last, st, st.height);
    }
  }
  refit_tree(*t, tets, X, Y, Z);
  return t;
}

// ------------------------------------------------------- tet-tet narrow phase
template <class T> struct Tet { V3<T> p[4]; };
// geometry_tetrahedron.h:231-257 (faces opposite vertex 0..3: 123, 032, 013, 021)
static const int FACE[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

template <class T> struct UPlane { V3<T> n, p; };  // contacts ...tetrahedron.h:20-28
// contacts ...tetrahedron.h:30-62 (three ifs, a later tie overwrites an earlier one)
template <class T> inline UPlane<T> make_unscaled_plane(const V3<T>& p0, const V3<T>& p1, const V3<T>& p2) {
  const V3<T> m0 = cross(sub(p1, p0), sub(p2, p0));
  const V3<T> m1 = cross(sub(p2, p1), sub(p0, p1));
  const V3<T> m2 = cross(sub(p0, p2), sub(p1, p2));
  const T l0 = dot(m0, m0), l1 = dot(m1, m1), l2 = dot(m2, m2);
  UPlane<T> pl{V3<T>{0, 0, 0}, V3<T>{0, 0, 0}};
  if (l0 >= l1 && l0 >= l2) { pl.n = m0; pl.p = p0; }
  if (l1 >= l0 && l1 >= l2) { pl.n = m1; pl.p = p1; }
  if (l2 >= l1 && l2 >= l0) { pl.n = m2; pl.p = p2; }
  return pl;
}
template <class T> inline UPlane<T> face_uplane(const Tet<T>& t, int f) {
  return make_unscaled_plane(t.p[FACE[f][0]], t.p[FACE[f][1]], t.p[FACE[f][2]]);
}
template <class T> inline T usd(const V3<T>& q, const UPlane<T>& pl) { return dot(pl.n, sub(q, pl.p)); }  // :81-85
template <class T> inline bool inside_planes(const V3<T>& q, const UPlane<T>* P) {                      // :87-110
  const T d0 = usd(q, P[0]), d1 = usd(q, P[1]), d2 = usd(q, P[2]), d3 = usd(q, P[3]);
  if (d0 > T(0)) return false;
  if (d1 > T(0)) return false;
  if (d2 > T(0)) return false;
  if (d3 > T(0)) return false;
  return true;
}
template <class T> inline bool is_crossing_plane(const V3<T>& a, const V3<T>& b, const UPlane<T>& pl) {  // :141-157
  const T dA = usd(a, pl), dB = usd(b, pl);
  if (dA >= T(0) && dB >= T(0)) return false;
  if (dA <= T(0) && dB <= T(0)) return false;
  return true;
}
template <class T> inline V3<T> make_intersection(const V3<T>& a, const V3<T>& b, const UPlane<T>& pl) {  // :112-125
  const T dA = std::fabs(usd(a, pl));
  const T dB = std::fabs(usd(b, pl));
  const T t = dA / (dA + dB);
  return add(a, mul(sub(b, a), t));
}

// types/geometry_plane.h:103-111 + :66-71 (the Plane constructor normalises the
// already-unit normal once more; the offset uses the once-normalised normal).
template <class T> struct Plane { V3<T> n; T w; };
template <class T> inline Plane<T> make_plane(const V3<T>& p0, const V3<T>& p1, const V3<T>& p2) {
  const V3<T> normal = unit(cross(sub(p1, p0), sub(p2, p0)));
  const T offset = dot(p0, normal);
  return Plane<T>{unit(normal), offset};
}
template <class T> inline T signed_distance(const V3<T>& p, const Plane<T>& P) { return dot(P.n, p) - P.w; }  // :171-175

// overlap/geometry_overlap_tetrahedron_tetrahedron.h:14-46
template <class T> inline bool SAT_line(const V3<T>& a0, const V3<T>& a1, const V3<T>& a2,
                                        const V3<T>& b0, const V3<T>& b1, const V3<T>& b2, const V3<T>& n) {
  const T t0_d0 = dot(n, a0), t0_d1 = dot(n, a1), t0_d2 = dot(n, a2);
  const T t1_d0 = dot(n, b0), t1_d1 = dot(n, b1), t1_d2 = dot(n, b2);
  const T t0_min = smin(smin(t0_d0, t0_d1), t0_d2);
  const T t1_min = smin(smin(t1_d0, t1_d1), t1_d2);
  const T t0_max = smax(smax(t0_d0, t0_d1), t0_d2);
  const T t1_max = smax(smax(t1_d0, t1_d1), t1_d2);
  return (t0_max < t1_min) || (t0_min > t1_max);
}
// Edge e = (E0[e], E1[e]) as vertex indices; the reference enumerates edges
// 01, 02, 03, 12, 13, 23 (overlap ...tetrahedron.h:102-107).
static const int EDGE[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
// Which three vertices SAT_line projects for each edge-edge axis: the reference
// hard-codes these triples (overlap ...tetrahedron.h:109-287).  A-side triple by
// A edge; B-side triple by (A edge, B edge).  Triples: 0={1,2,3} 1={0,1,3} 2={0,1,2}.
static const int TRIPLE[3][3] = {{1, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int SAT_A_TRIPLE[6] = {0, 0, 0, 1, 2, 2};
static const int SAT_B_TRIPLE[6][6] = {
    {0, 1, 2, 1, 2, 2}, {0, 0, 2, 1, 2, 2}, {0, 0, 0, 1, 2, 2},
    {0, 0, 0, 1, 2, 2}, {0, 0, 0, 1, 2, 2}, {0, 0, 0, 1, 2, 2}};

// overlap/geometry_overlap_tetrahedron_tetrahedron.h:50-290
template <class T> inline bool overlap_tet_tet(const Tet<T>& A, const Tet<T>& B) {
  for (int i = 0; i < 4; ++i) {
    const Plane<T> pl = make_plane(A.p[FACE[i][0]], A.p[FACE[i][1]], A.p[FACE[i][2]]);
    const T b0 = signed_distance(B.p[0], pl), b1 = signed_distance(B.p[1], pl);
    const T b2 = signed_distance(B.p[2], pl), b3 = signed_distance(B.p[3], pl);
    if (b0 > T(0) && b1 > T(0) && b2 > T(0) && b3 > T(0)) return false;
  }
  for (int i = 0; i < 4; ++i) {
    const Plane<T> pl = make_plane(B.p[FACE[i][0]], B.p[FACE[i][1]], B.p[FACE[i][2]]);
    const T a0 = signed_distance(A.p[0], pl), a1 = signed_distance(A.p[1], pl);
    const T a2 = signed_distance(A.p[2], pl), a3 = signed_distance(A.p[3], pl);
    if (a0 > T(0) && a1 > T(0) && a2 > T(0) && a3 > T(0)) return false;
  }
  for (int ea = 0; ea < 6; ++ea) {
    const V3<T> eA = sub(A.p[EDGE[ea][1]], A.p[EDGE[ea][0]]);
    const int* ta = TRIPLE[SAT_A_TRIPLE[ea]];
    for (int eb = 0; eb < 6; ++eb) {
      const V3<T> eB = sub(B.p[EDGE[eb][1]], B.p[EDGE[eb][0]]);
      const V3<T> n = cross(eA, eB);
      const int* tb = TRIPLE[SAT_B_TRIPLE[ea][eb]];
      if (SAT_line(A.p[ta[0]], A.p[ta[1]], A.p[ta[2]], B.p[tb[0]], B.p[tb[1]], B.p[tb[2]], n)) return false;
    }
  }
  return true;
}

// contacts ...tetrahedron.h:464-533.  Returns the chosen A face (or -1).
template <class T> inline int pick_most_opposing(const Tet<T>& A, const Tet<T>& B, const bool* sA, const bool* sB, V3<T>& n, int* face_b = nullptr) {
  int found = -1;
  T best_fit = T(0);
  for (int a = 0; a < 4; ++a) {
    if (!sA[a]) continue;
    const UPlane<T> plA = face_uplane(A, a);
    const V3<T> nA = unit(plA.n);
    for (int b = 0; b < 4; ++b) {
      if (!sB[b]) continue;
      const UPlane<T> plB = face_uplane(B, b);
      const V3<T> nB = unit(plB.n);
      const T test = dot(nA, nB);
      if (test >= T(0)) continue;
      const V3<T>& pB = B.p[b];
      const V3<T>& pA = A.p[a];
      const T testA = dot(plA.n, sub(pB, plA.p));
      const T testB = dot(plB.n, sub(pA, plB.p));
      if (testA < T(0)) continue;
      if (testB < T(0)) continue;
      if (test < best_fit) {
        n = nA;
        best_fit = test;
        found = a;
        if (face_b) *face_b = b;
      }
    }
  }
  return found;
}

// One candidate contact point with its provenance (what the adjoint freezes).
//   kind 0: vertex v of B inside A      kind 1: vertex v of A inside B
//   kind 2: edge e of B crossing plane f of A   kind 3: edge e of A crossing plane f of B
struct Cand { int kind, v, e, f; };
// One-byte provenance code, the same encoding the HIP narrow phase records:
// kind | (vertex or edge) << 2 | plane << 5.
inline uint8_t cand_code(const Cand& c) { return (uint8_t)(c.kind | ((c.kind < 2 ? c.v : c.e) << 2) | (c.f < 0 ? 0 : c.f << 5)); }

template <class T> struct PairContacts {
  int face_a = -1;               // A face whose unit normal is the contact normal
  V3<T> n{0, 0, 0};
  T depth = T(0);
  int imin = -1, imax = -1;      // candidates attaining min / max of p.n
  std::vector<Cand> cand;        // all candidates, in generation order
  std::vector<V3<T>> raw;        // candidate points before projection
  std::vector<V3<T>> proj;       // after projection onto the mid plane
  std::vector<uint8_t> unique;   // reported (1) or filtered as duplicate (0)
};

// contacts ...tetrahedron.h:543-681 (generate_contacts_from_intersection)
template <class T> inline bool generate_contacts(const Tet<T>& A, const Tet<T>& B, const V3<T>& n, PairContacts<T>& out) {
  UPlane<T> PA[4], PB[4];
  for (int v = 0; v < 4; ++v) { PA[v] = face_uplane(A, v); PB[v] = face_uplane(B, v); }
  out.cand.clear(); out.raw.clear();
  for (int v = 0; v < 4; ++v) {
    if (inside_planes(B.p[v], PA)) { out.cand.push_back(Cand{0, v, -1, -1}); out.raw.push_back(B.p[v]); }
    if (inside_planes(A.p[v], PB)) { out.cand.push_back(Cand{1, v, -1, -1}); out.raw.push_back(A.p[v]); }
  }
  for (int e = 0; e < 6; ++e) {
    const int i = EDGE[e][0], j = EDGE[e][1];
    for (int p = 0; p < 4; ++p) {
      if (is_crossing_plane(B.p[i], B.p[j], PA[p])) {
        const V3<T> qB = make_intersection(B.p[i], B.p[j], PA[p]);
        if (inside_planes(qB, PA)) { out.cand.push_back(Cand{2, -1, e, p}); out.raw.push_back(qB); }
      }
      if (is_crossing_plane(A.p[i], A.p[j], PB[p])) {
        const V3<T> qA = make_intersection(A.p[i], A.p[j], PB[p]);
        if (inside_planes(qA, PB)) { out.cand.push_back(Cand{3, -1, e, p}); out.raw.push_back(qA); }
      }
    }
  }
  if (out.raw.empty()) return false;
  T min_val = std::numeric_limits<T>::max();
  T max_val = -std::numeric_limits<T>::max();
  for (size_t c = 0; c < out.raw.size(); ++c) {
    const T d = dot(out.raw[c], n);
    if (d < min_val) out.imin = (int)c;          // first attaining index, as std::min keeps
    if (max_val < d) out.imax = (int)c;
    min_val = smin(min_val, d);
    max_val = smax(max_val, d);
  }
  out.depth = min_val - max_val;
  const V3<T> mid = mul(mul(n, max_val + min_val), T(0.5));
  out.proj.resize(out.raw.size());
  for (size_t c = 0; c < out.raw.size(); ++c) {
    const T s = dot(n, sub(out.raw[c], mid));
    out.proj[c] = sub(out.raw[c], V3<T>{s * n.x, s * n.y, s * n.z});
  }
  out.unique.assign(out.raw.size(), 1);
  const T wp = working_precision<T>();
  for (size_t c = 0; c < out.proj.size(); ++c)
    for (size_t q = 0; q < c; ++q)
      if (norm_1(sub(out.proj[q], out.proj[c])) < wp) { out.unique[c] = 0; break; }
  return true;
}

// contacts ...tetrahedron.h:1312-1345 (MOST_OPPOSING_SURFACES)
template <class T> inline bool contacts_most_opposing(const Tet<T>& A, const Tet<T>& B, const bool* sA, const bool* sB, PairContacts<T>& out) {
  out.face_a = -1;
  out.cand.clear(); out.raw.clear(); out.proj.clear(); out.unique.clear();
  if (!overlap_tet_tet(A, B)) return false;
  V3<T> n{0, 0, 0};
  const int fa = pick_most_opposing(A, B, sA, sB, n);
  if (fa < 0) return false;
  out.face_a = fa;
  out.n = n;
  return generate_contacts(A, B, n, out);
}

// ------------------------------------------- frozen-choice evaluation (FP64)
// The penetration measure of one tet pair with every discrete choice of the
// forward pass frozen: the A face whose unit normal is n, and the provenance
// of the candidates attaining min and max of p.n.  depth = (p_min - p_max).n
// (contacts ...tetrahedron.h:543-681 restricted to those two points).  This is
// what the adjoint differentiates; it is not part of the reference.
// Edge-plane crossing with the signed parameter t = sA / (sA - sB): equal to
// make_intersection's |sA| / (|sA| + |sB|) whenever the edge crosses the
// plane (the frozen choice), and smooth where the unsigned form has a kink.
template <class T> inline V3<T> crossing_frozen(const V3<T>& a, const V3<T>& b, const UPlane<T>& pl) {
  const T sA = usd(a, pl), sB = usd(b, pl);
  const T t = sA / (sA - sB);
  return add(a, mul(sub(b, a), t));
}
template <class T> inline V3<T> cand_point(const Tet<T>& A, const Tet<T>& B, uint8_t code) {
  const int kind = code & 3, ve = (code >> 2) & 7, f = (code >> 5) & 3;
  if (kind == 0) return B.p[ve];
  if (kind == 1) return A.p[ve];
  if (kind == 2) return crossing_frozen(B.p[EDGE[ve][0]], B.p[EDGE[ve][1]], face_uplane(A, f));
  return crossing_frozen(A.p[EDGE[ve][0]], A.p[EDGE[ve][1]], face_uplane(B, f));
}
template <class T> inline T frozen_depth(const Tet<T>& A, const Tet<T>& B, int face, uint8_t cmin, uint8_t cmax) {
  const V3<T> n = unit(face_uplane(A, face).n);
  return dot(cand_point(A, B, cmin), n) - dot(cand_point(A, B, cmax), n);
}

// Reverse mode of the frozen depth: adds d(depth)/d(vertex) * w to gA, gB.
// Plane through (p0,p1,p2) as N = (p1-p0) x (p2-p0), point p0 (the three
// choices make_unscaled_plane picks between are equal in exact arithmetic).
inline void grad_cross_plane(const V3<double>& p0, const V3<double>& p1, const V3<double>& p2, const V3<double>& gN,
                             V3<double>& g0, V3<double>& g1, V3<double>& g2) {
  const V3<double> e1 = sub(p1, p0), e2 = sub(p2, p0);
  const V3<double> ge1 = cross(e2, gN), ge2 = cross(gN, e1);
  g1 = add(g1, ge1); g2 = add(g2, ge2); g0 = sub(g0, add(ge1, ge2));
}
inline void grad_cand(const Tet<double>& A, const Tet<double>& B, uint8_t code, const V3<double>& gq, V3<double>* gA, V3<double>* gB) {
  const int kind = code & 3, ve = (code >> 2) & 7, f = (code >> 5) & 3;
  if (kind == 0) { gB[ve] = add(gB[ve], gq); return; }
  if (kind == 1) { gA[ve] = add(gA[ve], gq); return; }
  // edge (a,b) of one tet crossing the face plane f of the other
  const Tet<double>& E = (kind == 2) ? B : A;   // tet owning the edge
  const Tet<double>& F = (kind == 2) ? A : B;   // tet owning the plane
  V3<double>* gE = (kind == 2) ? gB : gA;
  V3<double>* gF = (kind == 2) ? gA : gB;
  const int ia = EDGE[ve][0], ib = EDGE[ve][1];
  const V3<double>& a = E.p[ia];
  const V3<double>& b = E.p[ib];
  const int f0 = FACE[f][0], f1 = FACE[f][1], f2 = FACE[f][2];
  const V3<double> N = cross(sub(F.p[f1], F.p[f0]), sub(F.p[f2], F.p[f0]));
  const double sA = dot(N, sub(a, F.p[f0])), sB = dot(N, sub(b, F.p[f0]));
  const double den = sA - sB;
  const double t = sA / den;
  gE[ia] = add(gE[ia], mul(gq, 1.0 - t));
  gE[ib] = add(gE[ib], mul(gq, t));
  const double gt = dot(sub(b, a), gq);
  const double gsA = gt * (-sB / (den * den)), gsB = gt * (sA / (den * den));
  gE[ia] = add(gE[ia], mul(N, gsA));
  gE[ib] = add(gE[ib], mul(N, gsB));
  const V3<double> gN = add(mul(sub(a, F.p[f0]), gsA), mul(sub(b, F.p[f0]), gsB));
  gF[f0] = sub(gF[f0], mul(N, gsA + gsB));
  grad_cross_plane(F.p[f0], F.p[f1], F.p[f2], gN, gF[f0], gF[f1], gF[f2]);
}
inline double frozen_depth_grad(const Tet<double>& A, const Tet<double>& B, int face, uint8_t cmin, uint8_t cmax, double w,
                                V3<double>* gA, V3<double>* gB) {
  const int f0 = FACE[face][0], f1 = FACE[face][1], f2 = FACE[face][2];
  const V3<double> N = cross(sub(A.p[f1], A.p[f0]), sub(A.p[f2], A.p[f0]));
  const double len = std::sqrt(dot(N, N));
  const V3<double> n = mul(N, 1.0 / len);
  const V3<double> pmin = cand_point(A, B, cmin), pmax = cand_point(A, B, cmax);
  const double depth = dot(pmin, n) - dot(pmax, n);
  grad_cand(A, B, cmin, mul(n, w), gA, gB);
  grad_cand(A, B, cmax, mul(n, -w), gA, gB);
  const V3<double> gn = mul(sub(pmin, pmax), w);
  const V3<double> gN = mul(sub(gn, mul(n, dot(n, gn))), 1.0 / len);
  grad_cross_plane(A.p[f0], A.p[f1], A.p[f2], gN, gA[f0], gA[f1], gA[f2]);
  return depth;
}
// Pose gradient of r = Im(q r0 q*) + p = (s^2 - u.u) r0 + 2 (u.r0) u + 2 s (u x r0) + p.
inline void pose_grad(const double q[4], const V3<double>& r0, const V3<double>& g, double* gp, double* gq) {
  const V3<double> u = mk(q[0], q[1], q[2]);
  const double s = q[3];
  gp[0] += g.x; gp[1] += g.y; gp[2] += g.z;
  const V3<double> gu = add(add(mul(u, -2.0 * dot(g, r0)), mul(r0, 2.0 * dot(g, u))), add(mul(g, 2.0 * dot(u, r0)), mul(cross(r0, g), 2.0 * s)));
  gq[0] += gu.x; gq[1] += gu.y; gq[2] += gu.z;
  gq[3] += dot(g, add(mul(r0, 2.0 * s), mul(cross(u, r0), 2.0)));
}

// -------------------------------------------------------------- traversal
struct Contact { uint32_t ta, tb; float p[3], n[3], d; };
struct Mesh {
  const uint32_t* tets; const float *X, *Y, *Z; const uint8_t* surf;  // surf: bit v = face opposite vertex v
};
struct Sink {
  uint64_t tests = 0, branch_tests = 0, descent_tests = 0, branch_overlaps = 0, p1 = 0, p2 = 0, ncontacts = 0, narrow_hits = 0;
  std::vector<uint32_t>* p1v = nullptr;  // (tetA, tetB) for every overlapping leaf pair
  std::vector<uint32_t>* p2v = nullptr;  // surface-filtered subset
  std::vector<Contact>* cv = nullptr;
  bool do_narrow = true;
  double t_exact = 0.0;
};

inline Tet<float> gather_tet(const Mesh& m, uint32_t t) {
  Tet<float> out;
  for (int v = 0; v < 4; ++v) { const uint32_t i = m.tets[4 * t + v]; out.p[v] = mk(m.X[i], m.Y[i], m.Z[i]); }
  return out;
}

// kdop_tandem_traversal.h:18-156
inline void traversal(size_t ia, const SubTree& BA, const Mesh& mA, size_t ib, const SubTree& BB, const Mesh& mB, int N, int depth, Sink& s) {
  const Node& nA = BA.nodes[ia];
  const Node& nB = BB.nodes[ib];
  ++s.tests;
  if (depth == 0) ++s.branch_tests; else ++s.descent_tests;
  if (!dop_overlap(nA.vol, nB.vol, N)) return;
  if (depth == 0) ++s.branch_overlaps;
  const bool la = nA.is_leaf(), lb = nB.is_leaf();
  if (la && lb) {
    const uint32_t ta = (uint32_t)nA.start, tb = (uint32_t)nB.start;
    ++s.p1;
    if (s.p1v) { s.p1v->push_back(ta); s.p1v->push_back(tb); }
    const uint8_t fa = mA.surf[ta], fb = mB.surf[tb];
    if (!(fa & 15u)) return;  // :61-72 all faces of A internal
    if (!(fb & 15u)) return;  // :74-88
    ++s.p2;
    if (s.p2v) { s.p2v->push_back(ta); s.p2v->push_back(tb); }
    if (!s.do_narrow) return;
    const Tet<float> A = gather_tet(mA, ta), B = gather_tet(mB, tb);
    bool sA[4], sB[4];
    for (int v = 0; v < 4; ++v) { sA[v] = (fa >> v) & 1u; sB[v] = (fb >> v) & 1u; }
    PairContacts<float> pc;
    if (contacts_most_opposing(A, B, sA, sB, pc)) {
      ++s.narrow_hits;
      for (size_t c = 0; c < pc.proj.size(); ++c) {
        if (!pc.unique[c]) continue;
        ++s.ncontacts;
        if (s.cv) {
          Contact ct{ta, tb, {pc.proj[c].x, pc.proj[c].y, pc.proj[c].z}, {pc.n.x, pc.n.y, pc.n.z}, pc.depth};
          s.cv->push_back(ct);
        }
      }
    }
  } else if (!la && !lb) {
    for (size_t a = nA.start; a <= nA.end; ++a)
      for (size_t b = nB.start; b <= nB.end; ++b) traversal(a, BA, mA, b, BB, mB, N, depth + 1, s);
  } else if (!la && lb) {
    for (size_t a = nA.start; a <= nA.end; ++a) traversal(a, BA, mA, ib, BB, mB, N, depth + 1, s);
  } else {
    for (size_t b = nB.start; b <= nB.end; ++b) traversal(ia, BA, mA, b, BB, mB, N, depth + 1, s);
  }
}

// kdop_tandem_traversal.h:160-196 over the branch rows [a0, a1)
inline void tandem_rows(const Tree& A, const Mesh& mA, const Tree& B, const Mesh& mB, size_t a0, size_t a1, Sink& s) {
  const auto& brA = A.levels.back();
  const auto& brB = B.levels.back();
  for (size_t a = a0; a < a1; ++a)
    for (size_t b = 0; b < brB.size(); ++b) traversal(0, brA[a], mA, 0, brB[b], mB, A.N, 0, s);
}

// kdop_mesh_reorder.h:172-353 (Morton order; the default of :364-387).  The
// reference's std::sort leaves equal codes in an implementation-defined
// order; here they keep their input order (a stable sort), the one change.
// Vertices that no tetrahedron names are appended in input order (the
// reference's output mesh drops them).
inline void mesh_reorder(uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt, const uint32_t* T, float* Xo, float* Yo,
                         float* Zo, uint32_t* To) {
  std::vector<float> cx(nt), cy(nt), cz(nt);
  float min_X = std::numeric_limits<float>::max(), min_Y = min_X, min_Z = min_X;
  float max_X = std::numeric_limits<float>::min(), max_Y = max_X, max_Z = max_X;  // :197-199 as written
  for (uint32_t i = 0; i < nt; ++i) {                                               // :201-231
    const uint32_t* t = T + 4 * (size_t)i;
    cx[i] = (float)(0.25 * (X[t[0]] + X[t[1]] + X[t[2]] + X[t[3]]));
    cy[i] = (float)(0.25 * (Y[t[0]] + Y[t[1]] + Y[t[2]] + Y[t[3]]));
    cz[i] = (float)(0.25 * (Z[t[0]] + Z[t[1]] + Z[t[2]] + Z[t[3]]));
    min_X = std::min(min_X, cx[i]); min_Y = std::min(min_Y, cy[i]); min_Z = std::min(min_Z, cz[i]);
    max_X = std::max(max_X, cx[i]); max_Y = std::max(max_Y, cy[i]); max_Z = std::max(max_Z, cz[i]);
  }
  const int KB = (int)(sizeof(unsigned long long) * 8) / 3;  // :235
  // :237-249: int scale = ((1 << K) - 1) / (max - min), the smallest over the
  // axes; a zero-width axis (undefined in the reference) counts as INT_MAX
  auto axis_scale = [&](float lo, float hi) {
    const float q = ((1 << KB) - 1) / (hi - lo);
    return q < 2147483647.0f ? (int)q : std::numeric_limits<int>::max();
  };
  int scale = axis_scale(min_X, max_X);
  scale = std::min(scale, axis_scale(min_Y, max_Y));
  scale = std::min(scale, axis_scale(min_Z, max_Z));
  std::vector<std::pair<unsigned long long, uint32_t>> codes(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    // :251-258: the integer coordinate is stored back into the float attribute
    const float qx = (float)static_cast<int>((cx[i] - min_X) * scale);
    const float qy = (float)static_cast<int>((cy[i] - min_Y) * scale);
    const float qz = (float)static_cast<int>((cz[i] - min_Z) * scale);
    const unsigned long long x = (unsigned long long)qx, y = (unsigned long long)qy, z = (unsigned long long)qz;
    unsigned long long code = 0;
    for (unsigned long k = 0; k < (unsigned long)KB; ++k) {  // :274-280
      code |= ((x >> k) & 1) << (3 * k);
      code |= ((y >> k) & 1) << (3 * k + 1);
      code |= ((z >> k) & 1) << (3 * k + 2);
    }
    codes[i] = std::make_pair(code, i);
  }
  std::stable_sort(codes.begin(), codes.end(),
                   [](const std::pair<unsigned long long, uint32_t>& a, const std::pair<unsigned long long, uint32_t>& b) { return a.first < b.first; });
  const uint32_t UNASSIGNED = 0xFFFFFFFFu;  // :300-350: vertices in first-touch order
  std::vector<uint32_t> lookup(nv, UNASSIGNED);
  uint32_t pushed = 0;
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t* t = T + 4 * (size_t)codes[i].second;
    for (int q = 0; q < 4; ++q) {
      if (lookup[t[q]] == UNASSIGNED) {
        lookup[t[q]] = pushed;
        Xo[pushed] = X[t[q]]; Yo[pushed] = Y[t[q]]; Zo[pushed] = Z[t[q]];
        ++pushed;
      }
      To[4 * (size_t)i + q] = lookup[t[q]];
    }
  }
  for (uint32_t v = 0; v < nv; ++v)
    if (lookup[v] == UNASSIGNED) {
      Xo[pushed] = X[v]; Yo[pushed] = Y[v]; Zo[pushed] = Z[v];
      ++pushed;
    }
}

// ------------------------------------------------------------------ raycast
// geometry_inside.h:19-40
inline bool outside_dop(const V3<float>& p, const Dop& d, const V3<float>* D, int N, float threshold) {
  for (int k = 0; k < N; ++k) {
    const float o = dot(D[k], p);
    if (o < d.lo[k] - threshold) return true;
    if (o > d.hi[k] + threshold) return true;
  }
  return false;
}
// raycast/geometry_raycast_dop.h:19-99 (VT::lowest/highest = -/+FLT_MAX,
// VT::infinity = FLT_MAX, tiny_constants.h:139-145)
inline bool raycast_dop(const V3<float>& p, const V3<float>& r, const Dop& dop, const V3<float>* D, int N, float threshold, V3<float>& point,
                        float& length) {
  const float FMAX = std::numeric_limits<float>::max();
  length = FMAX;
  point = mk(0.0f, 0.0f, 0.0f);
  float t_min = -FMAX, t_max = FMAX;
  for (int k = 0; k < N; ++k) {
    const V3<float> d0 = neg(D[k]);
    const float w0 = -dop.lo[k];
    const float f0 = dot(d0, r);
    const V3<float> d1 = D[k];
    const float w1 = dop.hi[k];
    const float f1 = dot(d1, r);
    if (f0 == 0.0f) continue;
    float t_near = (w0 - dot(d0, p)) / f0;
    float t_far = (w1 - dot(d1, p)) / f1;
    if (t_near > t_far) std::swap(t_near, t_far);
    t_min = (t_near > t_min) ? t_near : t_min;
    t_max = (t_far < t_max) ? t_far : t_max;
  }
  if (t_min > 0.0f && t_min <= t_max) {
    const V3<float> q = add(p, mul(r, t_min));
    if (outside_dop(q, dop, D, N, threshold)) return false;
    length = t_min;
    point = q;
    return true;
  }
  return false;
}
// raycast/geometry_raycast_tetrahedron.h:27-131 with the faces of
// get_opposite_face (types/geometry_tetrahedron.h:231-257)
inline bool raycast_tet(const V3<float>& o, const V3<float>& r, const Tet<float>& tet, const bool surf[4], V3<float>& hit, float& length) {
  static const int F[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  length = std::numeric_limits<float>::max();
  hit = mk(0.0f, 0.0f, 0.0f);
  bool status = false;
  V3<float> n[4];
  float w[4];
  for (int m = 0; m < 4; ++m) {
    const V3<float>& a = tet.p[F[m][0]];
    const V3<float>& b = tet.p[F[m][1]];
    const V3<float>& c = tet.p[F[m][2]];
    const V3<float> n0 = cross(sub(b, a), sub(c, a));
    const V3<float> n1 = cross(sub(c, b), sub(a, b));
    const V3<float> n2 = cross(sub(a, c), sub(b, c));
    const float w0 = dot(n0, a), w1 = dot(n1, b), w2 = dot(n2, c);
    const float t0 = dot(n0, n0), t1 = dot(n1, n1), t2 = dot(n2, n2);
    n[m] = mk(0.0f, 0.0f, 0.0f);
    w[m] = 0.0f;
    if (t0 >= t1 && t0 >= t2) { n[m] = n0; w[m] = w0; }
    else if (t1 >= t0 && t1 >= t2) { n[m] = n1; w[m] = w1; }
    else if (t2 >= t0 && t2 >= t1) { n[m] = n2; w[m] = w2; }
  }
  for (int m = 0; m < 4; ++m) {
    if (!surf[m]) continue;
    const float rn = dot(r, n[m]);
    const float on = dot(o, n[m]);
    if (rn >= 0.0f) continue;
    const float t = (w[m] - on) / rn;
    if (t < 0.0f) continue;
    const V3<float> q = add(o, mul(r, t));
    const int i = (m + 1) % 4, j = (m + 2) % 4, k = (m + 3) % 4;
    if (dot(n[i], q) - w[i] > 0.0f) continue;
    if (dot(n[j], q) - w[j] > 0.0f) continue;
    if (dot(n[k], q) - w[k] > 0.0f) continue;
    if (t < length) {
      length = t;
      hit = q;
      status = true;
    }
  }
  return status;
}
// kdop_raycast.h:19-99
inline void raycast_node(const V3<float>& o, const V3<float>& r, size_t idx, const SubTree& br, const Mesh& m, const V3<float>* D, int N,
                         V3<float>& hit_point, float& length) {
  const float threshold = 0.01f;
  const Node& node = br.nodes[idx];
  V3<float> p;
  float t;
  if (!raycast_dop(o, r, node.vol, D, N, threshold, p, t)) return;
  if (t >= length) return;
  if (node.is_leaf()) {
    const uint32_t tet = (uint32_t)node.start;
    const Tet<float> T = gather_tet(m, tet);
    bool surf[4];
    for (int v = 0; v < 4; ++v) surf[v] = (m.surf[tet] >> v) & 1u;
    V3<float> q;
    float s;
    if (raycast_tet(o, r, T, surf, q, s)) {
      hit_point = s < length ? q : hit_point;
      length = s < length ? s : length;
    }
  } else {
    for (size_t c = node.start; c <= node.end; ++c) raycast_node(o, r, c, br, m, D, N, hit_point, length);
  }
}
// kdop_raycast.h:102-155; the ray is made by geometry::make_ray (direction
// normalised, types/geometry_ray.h:39-42).
inline bool raycast(const Tree& t, const Mesh& m, const V3<float>& origin, const V3<float>& direction, V3<float>& hit_point, float& length) {
  const V3<float> r = unit(direction);
  hit_point = mk(0.0f, 0.0f, 0.0f);
  length = std::numeric_limits<float>::max();
  V3<float> p;
  float tt;
  if (!raycast_dop(origin, r, t.root, t.D.data(), t.N, 0.01f, p, tt)) return false;
  for (const SubTree& br : t.levels.back()) raycast_node(origin, r, 0, br, m, t.D.data(), t.N, hit_point, length);
  return length < std::numeric_limits<float>::max();
}

// ------------------------------------------------------------ single traversal
struct SPlane { V3<float> n; float w; };
// types/geometry_plane.h:103-111, :156-163, :172-175
inline SPlane make_plane3(const V3<float>& p0, const V3<float>& p1, const V3<float>& p2) {
  const V3<float> n = unit(cross(sub(p1, p0), sub(p2, p0)));
  return SPlane{n, dot(p0, n)};
}
inline SPlane make_plane_np(const V3<float>& n, const V3<float>& p) { return SPlane{n, dot(p, n)}; }
inline float signed_dist(const V3<float>& p, const SPlane& P) { return dot(P.n, p) - P.w; }
// closest_points/geometry_closest_point_on_plane.h:14-28
inline V3<float> closest_point_on_plane(const V3<float>& p, const SPlane& P) {
  return sub(p, mul(P.n, dot(P.n, sub(p, mul(P.n, P.w)))));
}
// closest_points/geometry_closest_point_on_line.h:16-31, line from points (types/geometry_line.h:29-32,55-70)
inline V3<float> closest_point_on_line(const V3<float>& p, const V3<float>& a, const V3<float>& b) {
  const V3<float> d = unit(sub(b, a));
  return add(a, mul(d, dot(d, sub(p, a))));
}
// geometry_inside.h:135-161 (test_face_plane = false)
inline bool inside_triangle_nf(const V3<float>& p, const V3<float>& t0, const V3<float>& t1, const V3<float>& t2) {
  const V3<float> t[3] = {t0, t1, t2};
  const SPlane pl = make_plane3(t0, t1, t2);
  for (int k = 0; k < 3; ++k) {
    const V3<float> p2 = add(t[(k + 1) % 3], pl.n);
    if (signed_dist(p, make_plane3(t[k], t[(k + 1) % 3], p2)) > 0.0f) return false;
  }
  return true;
}
struct SphereContact { uint32_t sphere, tet; float p[3], n[3], d; };
// contacts/geometry_contacts_sphere_tetrahedron.h:35-195
inline bool contacts_sphere_tet(const V3<float>& c, float r, const Tet<float>& B, const bool surf[4], bool flip, V3<float>& out_p,
                                V3<float>& out_n, float& out_d) {
  static const int F[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  float d[4];
  SPlane plane[4];
  for (int k = 0; k < 4; ++k) {
    plane[k] = make_plane3(B.p[F[k][0]], B.p[F[k][1]], B.p[F[k][2]]);
    d[k] = signed_dist(c, plane[k]);
    if (d[k] > r) return false;
  }
  if (d[0] < 0.0f && d[1] < 0.0f && d[2] < 0.0f && d[3] < 0.0f) {
    for (int k = 0; k < 4; ++k) {
      if (d[k] < d[(k + 1) % 4]) continue;
      if (d[k] < d[(k + 2) % 4]) continue;
      if (d[k] < d[(k + 3) % 4]) continue;
      out_n = flip ? plane[k].n : neg(plane[k].n);
      out_p = closest_point_on_plane(c, plane[k]);
      out_d = d[k];
      return true;
    }
  }
  for (int i = 0; i < 4; ++i) {
    if (!(surf[(i + 1) % 4] || surf[(i + 2) % 4] || surf[(i + 3) % 4])) continue;
    const SPlane v1 = make_plane_np(unit(sub(B.p[i], B.p[(i + 1) % 4])), B.p[i]);
    const SPlane v2 = make_plane_np(unit(sub(B.p[i], B.p[(i + 2) % 4])), B.p[i]);
    const SPlane v3 = make_plane_np(unit(sub(B.p[i], B.p[(i + 3) % 4])), B.p[i]);
    if (signed_dist(c, v1) >= 0.0f && signed_dist(c, v2) >= 0.0f && signed_dist(c, v3) >= 0.0f) {
      const V3<float> m = sub(c, B.p[i]);
      const float dist = std::sqrt(dot(m, m));
      if (dist > r) return false;
      out_n = flip ? unit(m) : neg(unit(m));
      out_p = B.p[i];
      out_d = dist - r;
      return true;
    }
  }
  static const int E[6][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 2, 0, 3}, {1, 3, 2, 0}, {2, 3, 0, 1}};
  for (int v = 0; v < 6; ++v) {
    const int i = E[v][0], j = E[v][1], k = E[v][2], m = E[v][3];
    if (!(surf[k] || surf[m])) continue;
    const V3<float> e = sub(B.p[j], B.p[i]);
    const V3<float> nk = unit(cross(e, plane[k].n));
    const V3<float> nm = unit(cross(neg(e), plane[m].n));
    if (signed_dist(c, make_plane_np(nk, B.p[i])) >= 0.0f && signed_dist(c, make_plane_np(nm, B.p[i])) >= 0.0f) {
      const V3<float> q = closest_point_on_line(c, B.p[i], B.p[j]);
      const V3<float> mm = sub(c, q);
      const float dist = std::sqrt(dot(mm, mm));
      if (dist > r) return false;
      out_n = flip ? unit(mm) : neg(unit(mm));
      out_p = q;
      out_d = dist - r;
      return true;
    }
  }
  for (int v = 0; v < 4; ++v) {
    if (!surf[v]) continue;
    const float dist = signed_dist(c, plane[v]);
    if (dist < 0.0f) continue;
    if (inside_triangle_nf(c, B.p[F[v][0]], B.p[F[v][1]], B.p[F[v][2]])) {
      out_n = flip ? plane[v].n : neg(plane[v].n);
      out_p = closest_point_on_plane(c, plane[v]);
      out_d = dist - r;
      return true;
    }
  }
  return false;
}
// kdop_single_traversal.h:41-110
inline void single_node(const V3<float>& c, float r, const Dop& sd, size_t idx, const SubTree& br, const Mesh& m, int N, bool flip,
                        uint32_t sphere, std::vector<SphereContact>& out) {
  const Node& node = br.nodes[idx];
  if (!dop_overlap(node.vol, sd, N)) return;
  if (node.is_leaf()) {
    const uint32_t tet = (uint32_t)node.start;
    const Tet<float> T = gather_tet(m, tet);
    bool surf[4];
    for (int v = 0; v < 4; ++v) surf[v] = (m.surf[tet] >> v) & 1u;
    V3<float> p, n;
    float d;
    if (contacts_sphere_tet(c, r, T, surf, flip, p, n, d)) out.push_back(SphereContact{sphere, tet, {p.x, p.y, p.z}, {n.x, n.y, n.z}, d});
  } else {
    for (size_t ch = node.start; ch <= node.end; ++ch) single_node(c, r, sd, ch, br, m, N, flip, sphere, out);
  }
}
// kdop_single_traversal.h:123-161 with geometry::convert of the sphere (geometry_convert.h:13-31)
inline void single_traversal(const Tree& t, const Mesh& m, const V3<float>& c, float r, bool flip, uint32_t sphere,
                             std::vector<SphereContact>& out) {
  Dop sd;
  for (int k = 0; k < t.N; ++k) {
    const float o = dot(t.D[k], c);
    sd.lo[k] = o - r;
    sd.hi[k] = o + r;
  }
  if (!dop_overlap(t.root, sd, t.N)) return;
  for (const SubTree& br : t.levels.back()) single_node(c, r, sd, 0, br, m, t.N, flip, sphere, out);
}

}  // namespace orc

// =================================================================== C ABI
using namespace orc;

extern "C" {

// Direction table for K (K/2 unit vectors, xyz), float or double.
int orc_dir_table_f32(int K, float* out) {
  V3<float> D[13];
  const int n = dir_table<float>(K, D);
  for (int i = 0; i < n; ++i) { out[3 * i] = D[i].x; out[3 * i + 1] = D[i].y; out[3 * i + 2] = D[i].z; }
  return n;
}

// Morton reorder of a mesh (kdop_mesh_reorder.h:172-353, stable on equal codes).
int orc_mesh_reorder(uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt, const uint32_t* tets, float* Xo, float* Yo,
                     float* Zo, uint32_t* To) {
  if (nv == 0 || nt == 0) return -1;
  mesh_reorder(nv, X, Y, Z, nt, tets, Xo, Yo, Zo, To);
  return 0;
}

// Chunking shape (kdop_make_tree.h:216-229): out = {node_bytes, N_perfect, L, C, H}.
int orc_tree_shape(int K, uint64_t mem_bytes, uint64_t M, uint64_t* out) {
  size_t nb, np, L, C, H;
  if (!tree_shape(K, (size_t)mem_bytes, (size_t)M, nb, np, L, C, H)) return -1;
  out[0] = nb; out[1] = np; out[2] = L; out[3] = C; out[4] = H;
  return 0;
}

void* orc_tree_build(int K, uint64_t mem_bytes, uint32_t nt, const uint32_t* tets, const float* X, const float* Y, const float* Z) {
  return (void*)make_tree(K, (size_t)mem_bytes, nt, tets, X, Y, Z);
}
void orc_tree_free(void* h) { delete (Tree*)h; }
void orc_tree_refit(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z) {
  refit_tree(*(Tree*)h, tets, X, Y, Z);
}
// Per-level chunk counts (levels[0] = top).  Returns the number of levels.
int orc_tree_levels(void* h, uint64_t* chunks_per_level) {
  Tree* t = (Tree*)h;
  for (size_t l = 0; l < t->levels.size(); ++l) chunks_per_level[l] = t->levels[l].size();
  return (int)t->levels.size();
}
void orc_tree_root(void* h, float* out) {
  Tree* t = (Tree*)h;
  for (int k = 0; k < t->N; ++k) { out[2 * k] = t->root.lo[k]; out[2 * k + 1] = t->root.hi[k]; }
}
// Number of defined nodes over all chunks and levels.
uint64_t orc_tree_count_nodes(void* h) {
  Tree* t = (Tree*)h;
  uint64_t n = 0;
  for (auto& lv : t->levels) for (auto& st : lv) for (auto& nd : st.nodes) if (!nd.is_undefined()) ++n;
  return n;
}
// Export every defined node: info[6*i..] = {level, chunk, slot, is_leaf, start, end};
// vol[K*i..] = (lo0, hi0, lo1, hi1, ...).  Also the covered tet range
// [tet_first, tet_last] (tet space) in range[2*i..].
static void node_tet_range(const Tree* t, size_t level, size_t chunk, size_t slot, uint64_t& f, uint64_t& l) {
  const Node& nd = t->levels[level][chunk].nodes[slot];
  if (nd.is_leaf()) {
    if (level + 1 == t->levels.size()) { f = l = nd.start; return; }
    const size_t child = nd.start;  // child chunk root at level+1
    node_tet_range(t, level + 1, child, 0, f, l);
    return;
  }
  uint64_t f0, l0, f1, l1;
  node_tet_range(t, level, chunk, nd.start, f0, l0);
  node_tet_range(t, level, chunk, nd.end, f1, l1);
  f = f0; l = l1;
}
uint64_t orc_tree_export(void* h, uint64_t* info, uint64_t* range, float* vol) {
  Tree* t = (Tree*)h;
  uint64_t i = 0;
  for (size_t l = 0; l < t->levels.size(); ++l)
    for (size_t c = 0; c < t->levels[l].size(); ++c)
      for (size_t s = 0; s < t->levels[l][c].nodes.size(); ++s) {
        const Node& nd = t->levels[l][c].nodes[s];
        if (nd.is_undefined()) continue;
        info[6 * i] = l; info[6 * i + 1] = c; info[6 * i + 2] = s;
        info[6 * i + 3] = nd.is_leaf(); info[6 * i + 4] = nd.start; info[6 * i + 5] = nd.end;
        node_tet_range(t, l, c, s, range[2 * i], range[2 * i + 1]);
        for (int k = 0; k < t->N; ++k) { vol[(size_t)t->K * i + 2 * k] = nd.vol.lo[k]; vol[(size_t)t->K * i + 2 * k + 1] = nd.vol.hi[k]; }
        ++i;
      }
  return i;
}

// narrow_update_kdop_bvh.h:154-165: r = rotate(q, r0) + p, q = (x, y, z, w).
void orc_transform_f32(uint32_t nv, const float* X0, const float* Y0, const float* Z0, const float* p, const float* q, float* X, float* Y, float* Z) {
  for (uint32_t v = 0; v < nv; ++v) {
    const V3<float> r = rotate<float>(q, mk(X0[v], Y0[v], Z0[v]));
    X[v] = r.x + p[0]; Y[v] = r.y + p[1]; Z[v] = r.z + p[2];
  }
}
void orc_transform_f64(uint32_t nv, const double* X0, const double* Y0, const double* Z0, const double* p, const double* q, double* X, double* Y, double* Z) {
  for (uint32_t v = 0; v < nv; ++v) {
    const V3<double> r = rotate<double>(q, mk(X0[v], Y0[v], Z0[v]));
    X[v] = r.x + p[0]; Y[v] = r.y + p[1]; Z[v] = r.z + p[2];
  }
}

// Leaf DOP of one tetrahedron (4 points xyz), K floats out.
void orc_make_dop_f32(int K, const float* pts, float* out) {
  V3<float> D[13];
  const int N = dir_table<float>(K, D);
  V3<float> P[4];
  for (int v = 0; v < 4; ++v) P[v] = mk(pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]);
  Dop d;
  make_dop(d, P, 4, D, N);
  for (int k = 0; k < N; ++k) { out[2 * k] = d.lo[k]; out[2 * k + 1] = d.hi[k]; }
}

// counters layout for orc_traverse: see Sink (8 x u64) + 3 phase times (double, seconds)
struct TravOut {
  uint64_t tests, branch_tests, descent_tests, branch_overlaps, p1, p2, ncontacts, narrow_hits;
};

// Sequential (threads<=1) or row-split multi-threaded tandem traversal of two
// trees (kdop_tandem_traversal.h:160-226).  Root test first; then every
// branch pair.  Outputs are written in the reference's traversal order (the
// threaded version concatenates per-row-block results in row order).
// Any of p1/p2/contacts may be null (counted only).  Capacities are in pairs /
// contacts; counts beyond capacity are reported but not written.
int orc_traverse(void* hA, int ntA_unused, const uint32_t* tetsA, const float* XA, const float* YA, const float* ZA, const uint8_t* surfA,
                 void* hB, int ntB_unused, const uint32_t* tetsB, const float* XB, const float* YB, const float* ZB, const uint8_t* surfB,
                 int do_narrow, int threads,
                 uint32_t* p1, uint64_t p1_cap, uint32_t* p2, uint64_t p2_cap, Contact* contacts, uint64_t c_cap,
                 uint64_t* counters, double* seconds) {
  (void)ntA_unused; (void)ntB_unused;
  const Tree& A = *(Tree*)hA;
  const Tree& B = *(Tree*)hB;
  if (A.K != B.K) return -1;
  const Mesh mA{tetsA, XA, YA, ZA, surfA}, mB{tetsB, XB, YB, ZB, surfB};
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(counters, 0, sizeof(uint64_t) * 9);
  counters[8] = 1;  // root-vs-root test (kdop_tandem_traversal.h:164-165)
  if (!dop_overlap(A.root, B.root, A.N)) {
    counters[0] = 1;
    if (seconds) *seconds = 0.0;
    return 0;
  }
  const size_t CA = A.levels.back().size();
  int T = threads < 1 ? 1 : threads;
  if ((size_t)T > CA) T = (int)CA;
  std::vector<Sink> sinks(T);
  std::vector<std::vector<uint32_t>> v1(T), v2(T);
  std::vector<std::vector<Contact>> vc(T);
  for (int i = 0; i < T; ++i) {
    sinks[i].do_narrow = do_narrow != 0;
    if (p1) sinks[i].p1v = &v1[i];
    if (p2) sinks[i].p2v = &v2[i];
    if (contacts) sinks[i].cv = &vc[i];
  }
  auto work = [&](int i) {
    const size_t a0 = CA * (size_t)i / T, a1 = CA * (size_t)(i + 1) / T;
    tandem_rows(A, mA, B, mB, a0, a1, sinks[i]);
  };
  if (T == 1) work(0);
  else {
    std::vector<std::thread> th;
    for (int i = 0; i < T; ++i) th.emplace_back(work, i);
    for (auto& x : th) x.join();
  }
  const auto t1 = std::chrono::steady_clock::now();
  Sink tot;
  uint64_t o1 = 0, o2 = 0, oc = 0;
  for (int i = 0; i < T; ++i) {
    const Sink& s = sinks[i];
    tot.tests += s.tests; tot.branch_tests += s.branch_tests; tot.descent_tests += s.descent_tests;
    tot.branch_overlaps += s.branch_overlaps; tot.p1 += s.p1; tot.p2 += s.p2; tot.ncontacts += s.ncontacts; tot.narrow_hits += s.narrow_hits;
    for (size_t j = 0; j < v1[i].size(); j += 2) { if (o1 < p1_cap) { p1[2 * o1] = v1[i][j]; p1[2 * o1 + 1] = v1[i][j + 1]; } ++o1; }
    for (size_t j = 0; j < v2[i].size(); j += 2) { if (o2 < p2_cap) { p2[2 * o2] = v2[i][j]; p2[2 * o2 + 1] = v2[i][j + 1]; } ++o2; }
    for (size_t j = 0; j < vc[i].size(); ++j) { if (oc < c_cap) contacts[oc] = vc[i][j]; ++oc; }
  }
  counters[0] = 1 + tot.tests;  // + root test
  counters[1] = tot.branch_tests;
  counters[2] = tot.descent_tests;
  counters[3] = tot.branch_overlaps;
  counters[4] = tot.p1;
  counters[5] = tot.p2;
  counters[6] = tot.ncontacts;
  counters[7] = tot.narrow_hits;
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

// Timing sample for the CPU baseline: the branch rows [row0, row1) of A
// against every branch of B, single threaded, narrow phase included, results
// counted but not stored.  counters as orc_traverse (without the root test).
int orc_traverse_rows(void* hA, const uint32_t* tetsA, const float* XA, const float* YA, const float* ZA, const uint8_t* surfA,
                      void* hB, const uint32_t* tetsB, const float* XB, const float* YB, const float* ZB, const uint8_t* surfB,
                      uint64_t row0, uint64_t row1, int do_narrow, uint64_t* counters, double* seconds) {
  const Tree& A = *(Tree*)hA;
  const Tree& B = *(Tree*)hB;
  if (A.K != B.K) return -1;
  const Mesh mA{tetsA, XA, YA, ZA, surfA}, mB{tetsB, XB, YB, ZB, surfB};
  const size_t CA = A.levels.back().size();
  if (row1 > CA) row1 = CA;
  Sink s;
  s.do_narrow = do_narrow != 0;
  const auto t0 = std::chrono::steady_clock::now();
  tandem_rows(A, mA, B, mB, (size_t)row0, (size_t)row1, s);
  const auto t1 = std::chrono::steady_clock::now();
  counters[0] = s.tests; counters[1] = s.branch_tests; counters[2] = s.descent_tests; counters[3] = s.branch_overlaps;
  counters[4] = s.p1; counters[5] = s.p2; counters[6] = s.ncontacts; counters[7] = s.narrow_hits; counters[8] = 0;
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

// All-cores timing sample: the branch rows [row0, row1) of A against every
// branch of B on `threads` threads that take rows one at a time (rows differ
// a lot in work), narrow phase included, results counted only.
int orc_traverse_rows_mt(void* hA, const uint32_t* tetsA, const float* XA, const float* YA, const float* ZA, const uint8_t* surfA,
                         void* hB, const uint32_t* tetsB, const float* XB, const float* YB, const float* ZB, const uint8_t* surfB,
                         uint64_t row0, uint64_t row1, int do_narrow, int threads, uint64_t* counters, double* seconds) {
  const Tree& A = *(Tree*)hA;
  const Tree& B = *(Tree*)hB;
  if (A.K != B.K) return -1;
  const Mesh mA{tetsA, XA, YA, ZA, surfA}, mB{tetsB, XB, YB, ZB, surfB};
  const size_t CA = A.levels.back().size();
  if (row1 > CA) row1 = CA;
  const int T = threads < 1 ? 1 : threads;
  std::vector<Sink> sinks(T);
  std::atomic<uint64_t> next(row0);
  auto work = [&](int i) {
    sinks[i].do_narrow = do_narrow != 0;
    for (;;) {
      const uint64_t r = next.fetch_add(1);
      if (r >= row1) break;
      tandem_rows(A, mA, B, mB, (size_t)r, (size_t)r + 1, sinks[i]);
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int i = 0; i < T; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
  const auto t1 = std::chrono::steady_clock::now();
  std::memset(counters, 0, sizeof(uint64_t) * 9);
  for (const Sink& s : sinks) {
    counters[0] += s.tests; counters[1] += s.branch_tests; counters[2] += s.descent_tests; counters[3] += s.branch_overlaps;
    counters[4] += s.p1; counters[5] += s.p2; counters[6] += s.ncontacts; counters[7] += s.narrow_hits;
  }
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

// kdop::raycast for n rays against one refitted tree (world coordinates X, Y, Z).
int orc_raycast(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z, const uint8_t* surf, uint32_t n,
                const float* origins, const float* directions, float* hits, float* lengths, uint8_t* did_hit) {
  const Tree& t = *(Tree*)h;
  const Mesh m{tets, X, Y, Z, surf};
  for (uint32_t i = 0; i < n; ++i) {
    V3<float> hp;
    float len;
    const bool hit = raycast(t, m, mk(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]),
                             mk(directions[3 * i], directions[3 * i + 1], directions[3 * i + 2]), hp, len);
    hits[3 * i] = hp.x; hits[3 * i + 1] = hp.y; hits[3 * i + 2] = hp.z;
    lengths[i] = len;
    did_hit[i] = hit ? 1 : 0;
  }
  return 0;
}

// kdop::single_traversal for n spheres (x, y, z, radius each) against one
// refitted tree; contacts of sphere 0 first, each in report order.  Returns
// the count (records beyond cap are not written).
uint64_t orc_single_traversal(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z, const uint8_t* surf, uint32_t n,
                              const float* spheres, int flip, SphereContact* out, uint64_t cap) {
  const Tree& t = *(Tree*)h;
  const Mesh m{tets, X, Y, Z, surf};
  std::vector<SphereContact> all;
  for (uint32_t i = 0; i < n; ++i)
    single_traversal(t, m, mk(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2]), spheres[4 * i + 3], flip != 0, i, all);
  for (size_t i = 0; i < all.size() && i < cap; ++i) out[i] = all[i];
  return all.size();
}

// PROX contact reduction (SIMULATION/PROX/PROX/include/prox_collision_detection.h:
// 340-345): std::sort of prox::ContactPoint, then std::unique.  The record
// is (body i, body j, p[3], n[3], depth); bodies are compared as numbers (the
// reference compares body pointers).  Writes the input positions of the kept
// contacts, in their sorted order, to out; returns their count.  The
// reference's std::sort is unstable; std::stable_sort here leaves contacts that
// are equal under operator< (equal in every compared field, up to the sign of
// a zero) in input order.
struct ProxPoint {
  uint32_t bi, bj;
  float p[3], n[3], d;
  uint32_t pos;
};
// ContactPoint::operator< (prox_contact_point.h:140-190)
static bool prox_less(const ProxPoint& a, const ProxPoint& b) {
  if (a.bi < b.bi) return true;
  if (a.bi > b.bi) return false;
  if (a.bj < b.bj) return true;
  if (a.bj > b.bj) return false;
  const float av[7] = {a.p[1], a.p[0], a.p[2], a.n[1], a.n[0], a.n[2], a.d};
  const float bv[7] = {b.p[1], b.p[0], b.p[2], b.n[1], b.n[0], b.n[2], b.d};
  for (int i = 0; i < 7; ++i) {
    if (av[i] < bv[i]) return true;
    if (av[i] > bv[i]) return false;
  }
  return false;
}
// ContactPoint::operator== (prox_contact_point.h:91-138) with
// tiny::working_precision<float>() = epsilon * 10 (tiny_precision.h:21-24);
// `self` is the contact std::unique kept last, `pt` the candidate.
static bool prox_equal(const ProxPoint& self, const ProxPoint& pt) {
  const float precision = std::numeric_limits<float>::epsilon() * 10;
  if (self.bi != pt.bi) return false;
  if (self.bj != pt.bj) return false;
  if (std::fabs(pt.p[1] - self.p[1]) > precision) return false;
  if (std::fabs(pt.p[0] - self.p[0]) > precision) return false;
  if (std::fabs(pt.p[2] - self.p[2]) > precision) return false;
  if (std::fabs(pt.n[1] - self.n[1]) > precision) return false;
  if (std::fabs(pt.n[0] - self.n[0]) > precision) return false;
  if (std::fabs(pt.n[2] - self.n[2]) > precision) return false;
  if (std::fabs(pt.d - self.d) > precision) return false;
  return true;
}
uint64_t orc_prox_reduce(uint64_t n, const uint32_t* bi, const uint32_t* bj, const float* p, const float* nrm, const float* d, uint32_t* out) {
  std::vector<ProxPoint> v(n);
  for (uint64_t i = 0; i < n; ++i) {
    ProxPoint& x = v[i];
    x.bi = bi[i];
    x.bj = bj[i];
    for (int k = 0; k < 3; ++k) { x.p[k] = p[3 * i + k]; x.n[k] = nrm[3 * i + k]; }
    x.d = d[i];
    x.pos = (uint32_t)i;
  }
  std::stable_sort(v.begin(), v.end(), prox_less);
  auto last = std::unique(v.begin(), v.end(), prox_equal);
  v.erase(last, v.end());
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].pos;
  return v.size();
}

// Brute force: every (leaf of A, leaf of B) whose leaf DOPs overlap.  The leaf
// DOP of tet t comes from the tree (so it is the refit value).  Independent of
// the tree traversal; pins the traversal's pair set.
uint64_t orc_brute_pairs(void* hA, void* hB, uint32_t* out, uint64_t cap) {
  const Tree& A = *(Tree*)hA;
  const Tree& B = *(Tree*)hB;
  std::vector<const Dop*> la, lb;
  std::vector<uint32_t> ta, tb;
  for (auto& st : A.levels.back()) for (auto& nd : st.nodes) if (!nd.is_undefined() && nd.is_leaf()) { la.push_back(&nd.vol); ta.push_back((uint32_t)nd.start); }
  for (auto& st : B.levels.back()) for (auto& nd : st.nodes) if (!nd.is_undefined() && nd.is_leaf()) { lb.push_back(&nd.vol); tb.push_back((uint32_t)nd.start); }
  uint64_t n = 0;
  for (size_t i = 0; i < la.size(); ++i)
    for (size_t j = 0; j < lb.size(); ++j)
      if (dop_overlap(*la[i], *lb[j], A.N)) { if (n < cap) { out[2 * n] = ta[i]; out[2 * n + 1] = tb[j]; } ++n; }
  return n;
}

// Narrow phase for one tetrahedron pair.  pa, pb: 12 values (4 vertices xyz).
// sa, sb: surface bits.  Outputs (up to 56 candidates):
//   n_out[3], depth_out[1], face_out[1] (A face or -1), pts_out[3*cap] (unique
//   projected points, in report order).  Returns the number of reported contacts
//   (0 when the pair is rejected).
}  // extern "C"

template <class T>
static int contacts_pair(const T* pa, const T* pb, int sa_bits, int sb_bits, T* n_out, T* depth_out, int* face_out, T* pts_out, int cap, int* ncand_out,
                         uint8_t* codes = nullptr, uint8_t* uniq = nullptr, int* iminmax = nullptr) {
  Tet<T> A, B;
  for (int v = 0; v < 4; ++v) { A.p[v] = mk(pa[3 * v], pa[3 * v + 1], pa[3 * v + 2]); B.p[v] = mk(pb[3 * v], pb[3 * v + 1], pb[3 * v + 2]); }
  bool sA[4], sB[4];
  for (int v = 0; v < 4; ++v) { sA[v] = (sa_bits >> v) & 1; sB[v] = (sb_bits >> v) & 1; }
  PairContacts<T> pc;
  *face_out = -1;
  if (ncand_out) *ncand_out = 0;
  if (!contacts_most_opposing(A, B, sA, sB, pc)) { *face_out = pc.face_a; return 0; }
  *face_out = pc.face_a;
  n_out[0] = pc.n.x; n_out[1] = pc.n.y; n_out[2] = pc.n.z;
  *depth_out = pc.depth;
  if (ncand_out) *ncand_out = (int)pc.raw.size();
  for (size_t c = 0; c < pc.cand.size() && c < 56; ++c) {
    if (codes) codes[c] = cand_code(pc.cand[c]);
    if (uniq) uniq[c] = pc.unique[c];
  }
  if (iminmax) { iminmax[0] = pc.imin; iminmax[1] = pc.imax; }
  int k = 0;
  for (size_t c = 0; c < pc.proj.size(); ++c) {
    if (!pc.unique[c]) continue;
    if (k < cap) { pts_out[3 * k] = pc.proj[c].x; pts_out[3 * k + 1] = pc.proj[c].y; pts_out[3 * k + 2] = pc.proj[c].z; }
    ++k;
  }
  return k;
}
extern "C" {

int orc_contacts_pair_f32(const float* pa, const float* pb, int sa, int sb, float* n, float* depth, int* face, float* pts, int cap, int* ncand) {
  return contacts_pair<float>(pa, pb, sa, sb, n, depth, face, pts, cap, ncand);
}
// Same, also returning the provenance code and report flag of every candidate
// and the indices of the candidates attaining min / max of p.n.
int orc_contacts_pair_codes_f32(const float* pa, const float* pb, int sa, int sb, float* n, float* depth, int* face, float* pts, int cap,
                                int* ncand, uint8_t* codes, uint8_t* uniq, int* iminmax) {
  return contacts_pair<float>(pa, pb, sa, sb, n, depth, face, pts, cap, ncand, codes, uniq, iminmax);
}

// Frozen-choice FP64 geometry of one pair (see frozen_depth): unit normal,
// depth, and the projected points of the reported candidates.
int orc_frozen_pair_f64(const double* pa, const double* pb, int face, int ncand, const uint8_t* codes, const uint8_t* uniq,
                        int imin, int imax, double* n_out, double* depth_out, double* pts_out) {
  Tet<double> A, B;
  for (int v = 0; v < 4; ++v) { A.p[v] = mk(pa[3 * v], pa[3 * v + 1], pa[3 * v + 2]); B.p[v] = mk(pb[3 * v], pb[3 * v + 1], pb[3 * v + 2]); }
  const V3<double> n = unit(face_uplane(A, face).n);
  const V3<double> pmin = cand_point(A, B, codes[imin]), pmax = cand_point(A, B, codes[imax]);
  const double mn = dot(pmin, n), mx = dot(pmax, n);
  n_out[0] = n.x; n_out[1] = n.y; n_out[2] = n.z;
  *depth_out = mn - mx;
  const V3<double> mid = mul(mul(n, mx + mn), 0.5);
  int k = 0;
  for (int c = 0; c < ncand; ++c) {
    if (!uniq[c]) continue;
    const V3<double> p = cand_point(A, B, codes[c]);
    const double s = dot(n, sub(p, mid));
    const V3<double> q = sub(p, mul(n, s));
    pts_out[3 * k] = q.x; pts_out[3 * k + 1] = q.y; pts_out[3 * k + 2] = q.z;
    ++k;
  }
  return k;
}

// Penetration measure of one body pair, Phi = sum over narrow records of
// nuniq * depth, evaluated in FP64 from the material coordinates and the
// poses with every discrete choice frozen (records: ta, tb, face, nuniq,
// cmin, cmax); with grads non-null also dPhi/d(pose) by reverse mode:
// grads = {gpA[3], gqA[4], gpB[3], gqB[4]} (q as x, y, z, w).
double orc_phi_frozen_f64(uint32_t nrec, const uint32_t* ta, const uint32_t* tb, const uint8_t* face, const uint8_t* nuniq,
                          const uint8_t* cmin, const uint8_t* cmax,
                          const uint32_t* tetsA, const double* XA, const double* YA, const double* ZA, const double* pA, const double* qA,
                          const uint32_t* tetsB, const double* XB, const double* YB, const double* ZB, const double* pB, const double* qB,
                          double* grads) {
  double phi = 0.0;
  if (grads) for (int i = 0; i < 14; ++i) grads[i] = 0.0;
  for (uint32_t r = 0; r < nrec; ++r) {
    Tet<double> A, B;
    V3<double> A0[4], B0[4];
    for (int v = 0; v < 4; ++v) {
      const uint32_t ia = tetsA[4 * ta[r] + v], ib = tetsB[4 * tb[r] + v];
      A0[v] = mk(XA[ia], YA[ia], ZA[ia]);
      B0[v] = mk(XB[ib], YB[ib], ZB[ib]);
      A.p[v] = add(rotate<double>(qA, A0[v]), mk(pA[0], pA[1], pA[2]));
      B.p[v] = add(rotate<double>(qB, B0[v]), mk(pB[0], pB[1], pB[2]));
    }
    const double w = (double)nuniq[r];
    if (!grads) { phi += w * frozen_depth(A, B, face[r], cmin[r], cmax[r]); continue; }
    V3<double> gA[4], gB[4];
    for (int v = 0; v < 4; ++v) { gA[v] = mk(0.0, 0.0, 0.0); gB[v] = mk(0.0, 0.0, 0.0); }
    phi += w * frozen_depth_grad(A, B, face[r], cmin[r], cmax[r], w, gA, gB);
    for (int v = 0; v < 4; ++v) {
      pose_grad(qA, A0[v], gA[v], grads, grads + 3);
      pose_grad(qB, B0[v], gB[v], grads + 7, grads + 10);
    }
  }
  return phi;
}
int orc_contacts_pair_f64(const double* pa, const double* pb, int sa, int sb, double* n, double* depth, int* face, double* pts, int cap, int* ncand) {
  return contacts_pair<double>(pa, pb, sa, sb, n, depth, face, pts, cap, ncand);
}
int orc_overlap_tet_tet_f32(const float* pa, const float* pb) {
  Tet<float> A, B;
  for (int v = 0; v < 4; ++v) { A.p[v] = mk(pa[3 * v], pa[3 * v + 1], pa[3 * v + 2]); B.p[v] = mk(pb[3 * v], pb[3 * v + 1], pb[3 * v + 2]); }
  return overlap_tet_tet(A, B) ? 1 : 0;
}

}  // extern "C"
