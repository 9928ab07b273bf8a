/* kdop_amd.hpp -- the reference's KDOP call shapes in C++ over the C ABI of
 * kdop_amd.h (header only).
 *
 * The reference drives its KDOP path from three call sites:
 *   narrow::make_kdop_bvh         -> kdop::make_tree      (kdop_make_tree.h:190-288)
 *   narrow::update_kdop_bvh(pool) -> transform + refit per body (narrow_update_kdop_bvh.h:126-176)
 *   dispatch_tetramesh_tetramesh  -> kdop::tandem_traversal(std::vector<TestPair>&)
 *                                    (kdop_tandem_traversal.h:198-226), contacts
 *                                    reported through each pair's ContactsCallback
 * This header gives the same three operations for a caller that keeps its own
 * mesh arrays: trees stay in device memory across steps, a whole pool is
 * refitted in one call, and a whole pair list is traversed in one call, after
 * which every pair's callback receives its contacts in the order the
 * sequential reference reports them (kdop_sort_contacts, KDOP_ORDER_SEQUENTIAL).
 * reduce_contacts then gives what PROX's contact reduction makes of those
 * contacts (prox_collision_detection.h:340-345), computed on the device.
 * raycast and single_traversal are kdop::raycast (kdop_raycast.h:106-155) and
 * kdop::single_traversal (kdop_single_traversal.h:123-161) on a tree in its
 * last refitted pose, one query as the reference takes it or many at once.
 * Errors throw kdop_amd::Error (the reference exits the process instead).
 * One Context per host thread.
 */
#ifndef KDOP_AMD_HPP
#define KDOP_AMD_HPP

#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "kdop_amd.h"

namespace kdop_amd {

struct Error : std::runtime_error {
  int code;
  Error(int rc, const std::string& what) : std::runtime_error(what), code(rc) {}
};

/* The tree of one body: kdop::Tree<T,K> (kdop_tree.h:12-200) is held by value
 * in narrow::Object; here the context holds it and the caller keeps the id. */
struct Tree {
  uint32_t body = 0xFFFFFFFFu;
};

/* geometry::ContactsCallback (contacts/geometry_contacts_callback.h:7-24). */
struct ContactsCallback {
  virtual ~ContactsCallback() {}
  virtual void operator()(const float p[3], const float n[3], float depth) = 0;
};

/* kdop::TestPair (kdop_test_pair.h:15-87): two trees and the pair's callback
 * (the meshes and world coordinates live with the trees here). */
struct TestPair {
  const Tree* tree_a;
  const Tree* tree_b;
  ContactsCallback* callback;
};

/* narrow::KDopBvhUpdateWorkItem (narrow_update_kdop_bvh.h:28-120): the body
 * and its pose, r = rotate(q, r0) + p with q = (x, y, z, w). */
struct UpdateWorkItem {
  const Tree* tree;
  float p[3];
  float q[4];
};

class Context {
 public:
  explicit Context(int K, uint64_t chunk_bytes = 8000, int device = 0) {
    kdop_config cfg;
    kdop_default_config(&cfg);
    cfg.K = K;
    cfg.chunk_bytes = chunk_bytes;
    cfg.device = device;
    const int rc = kdop_create(&c_, &cfg);
    if (rc != KDOP_OK) throw Error(rc, "kdop_create failed");
  }
  ~Context() { kdop_destroy(c_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  kdop_ctx* get() const { return c_; }

  /* kdop::make_tree(mem_bytes, ...) on the material coordinates: tets holds 4
   * vertex indices per tetrahedron, surface one byte per tetrahedron (bit v:
   * the face opposite its v-th vertex has no neighbour, compute_surface_map's
   * flags).  mem_bytes 0: the context's chunk_bytes. */
  Tree make_tree(const std::vector<float>& X, const std::vector<float>& Y, const std::vector<float>& Z, const std::vector<uint32_t>& tets,
                 const std::vector<uint8_t>& surface, uint64_t mem_bytes = 0) {
    if (X.size() != Y.size() || X.size() != Z.size() || tets.size() % 4 != 0 || surface.size() != tets.size() / 4)
      throw Error(KDOP_ERR_ARG, "make_tree: inconsistent mesh arrays");
    Tree t;
    const uint32_t nv = (uint32_t)X.size(), nt = (uint32_t)(tets.size() / 4);
    if (mem_bytes)
      check(kdop_add_body_ex(c_, mem_bytes, nv, X.data(), Y.data(), Z.data(), nt, tets.data(), surface.data(), &t.body));
    else
      check(kdop_add_body(c_, nv, X.data(), Y.data(), Z.data(), nt, tets.data(), surface.data(), &t.body));
    return t;
  }

  /* narrow::update_kdop_bvh(pool): every body of the pool transformed and
   * refitted in one call. */
  void update_kdop_bvh(const std::vector<UpdateWorkItem>& pool) {
    std::vector<uint32_t> bodies;
    std::vector<float> pos, quat;
    bodies.reserve(pool.size());
    for (const UpdateWorkItem& w : pool) {
      bodies.push_back(w.tree->body);
      pos.insert(pos.end(), w.p, w.p + 3);
      quat.insert(quat.end(), w.q, w.q + 4);
    }
    check(kdop_refit(c_, (uint32_t)bodies.size(), bodies.data(), pos.data(), quat.data()));
  }

  /* kdop::tandem_traversal(pool) with MOST_OPPOSING_SURFACES contacts: one
   * traversal of every pair, then each pair's callback is called with its
   * contacts in `order` (KDOP_ORDER_SEQUENTIAL: the sequential reference's
   * order; KDOP_ORDER_OPENCL: the OpenCL path's).  Returns the counters. */
  kdop_counters tandem_traversal(const std::vector<TestPair>& pool, int order = KDOP_ORDER_SEQUENTIAL) {
    std::vector<uint32_t> ids;
    ids.reserve(2 * pool.size());
    for (const TestPair& tp : pool) {
      ids.push_back(tp.tree_a->body);
      ids.push_back(tp.tree_b->body);
    }
    kdop_counters cnt;
    check(kdop_traverse(c_, (uint32_t)pool.size(), ids.data(), KDOP_NARROW, &cnt));
    pair_bodies_ = ids;
    check(kdop_sort_contacts(c_, order));
    uint64_t n = 0;
    check(kdop_get_contacts(c_, nullptr, 0, 0, &n));
    contacts_.resize((size_t)n);
    if (n) check(kdop_get_contacts(c_, contacts_.data(), n, 0, &n));
    for (const kdop_contact& k : contacts_) {
      ContactsCallback* cb = pool[k.pair].callback;
      if (cb) (*cb)(k.p, k.n, k.depth);
    }
    return cnt;
  }

  /* kdop::raycast: true and the nearest front-facing surface hit (hit_point,
   * length) when the ray from p along r (normalised here, as make_ray does)
   * hits the tree's mesh; else false, (0, 0, 0) and FLT_MAX. */
  bool raycast(const Tree& tree, const float p[3], const float r[3], float hit_point[3], float& length) {
    uint8_t hit = 0;
    check(kdop_raycast(c_, tree.body, 1, p, r, hit_point, &length, &hit, 0));
    return hit != 0;
  }
  /* n rays at once (3 floats per origin, direction and hit point). */
  void raycast(const Tree& tree, uint32_t n, const float* origins, const float* directions, float* hit_points, float* lengths,
               uint8_t* hits) {
    check(kdop_raycast(c_, tree.body, n, origins, directions, hit_points, lengths, hits, 0));
  }

  /* kdop::single_traversal with a sphere: the callback receives the sphere's
   * contacts with the tree's mesh in the reference's order. */
  void single_traversal(const float center[3], float radius, const Tree& tree, ContactsCallback& callback, bool should_flip = false) {
    const float sphere[4] = {center[0], center[1], center[2], radius};
    for (const kdop_sphere_contact& k : spheres_(tree, 1, sphere, should_flip)) callback(k.p, k.n, k.depth);
  }
  /* n spheres at once (x, y, z, radius each): the contacts grouped by sphere,
   * each sphere's in the reference's order. */
  std::vector<kdop_sphere_contact> single_traversal(const Tree& tree, uint32_t n, const float* spheres, bool should_flip = false) {
    return spheres_(tree, n, spheres, should_flip);
  }

  /* PROX contact reduction of the last tandem_traversal's contacts (the
   * std::sort + std::unique of prox_collision_detection.h:340-345 over the
   * ContactPoints its callbacks would have collected, body i and body j being
   * each pair's trees): sorted by (body i, body j, p.y, p.x, p.z, n.y, n.x,
   * n.z, depth), contacts within the working precision of the last one kept
   * dropped.  Bodies are the Tree::body numbers. */
  std::vector<kdop_prox_contact> reduce_contacts() {
    uint64_t n = 0;
    check(kdop_reduce_contacts(c_, &n));
    contacts_.resize((size_t)n);
    if (n) check(kdop_get_contacts(c_, contacts_.data(), n, 0, &n));
    std::vector<kdop_prox_contact> out((size_t)n);
    for (size_t i = 0; i < out.size(); ++i) {
      const kdop_contact& k = contacts_[i];
      kdop_prox_contact& o = out[i];
      o.body_i = pair_bodies_[2 * (size_t)k.pair];
      o.body_j = pair_bodies_[2 * (size_t)k.pair + 1];
      for (int j = 0; j < 3; ++j) {
        o.p[j] = k.p[j];
        o.n[j] = k.n[j];
      }
      o.depth = k.depth;
    }
    return out;
  }

 private:
  std::vector<kdop_sphere_contact> spheres_(const Tree& tree, uint32_t n, const float* spheres, bool flip) {
    std::vector<kdop_sphere_contact> out(std::max<size_t>(1, (size_t)n * 16));
    uint64_t cnt = 0;
    int rc = kdop_single_traversal(c_, tree.body, n, spheres, flip ? 1 : 0, out.data(), out.size(), &cnt, 0);
    if (rc == KDOP_ERR_CAPACITY) {
      out.resize((size_t)cnt);
      rc = kdop_single_traversal(c_, tree.body, n, spheres, flip ? 1 : 0, out.data(), out.size(), &cnt, 0);
    }
    check(rc);
    out.resize((size_t)cnt);
    return out;
  }
  kdop_ctx* c_ = nullptr;
  std::vector<kdop_contact> contacts_;
  std::vector<uint32_t> pair_bodies_;  // body ids of the last traversal's pairs
  void check(int rc) const {
    if (rc != KDOP_OK) throw Error(rc, kdop_last_error(c_));
  }
};

}  // namespace kdop_amd

#endif  // KDOP_AMD_HPP
