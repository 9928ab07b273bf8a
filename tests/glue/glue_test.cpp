// C++ caller of include/kdop_amd.hpp against the oracle (test infrastructure):
// three box meshes in the reference's call shapes (make_tree, a pool refit,
// a pair-list tandem traversal with one ContactsCallback per pair), two pose
// steps; every pair's callback must receive exactly the contacts the oracle
// reports for that pair, in the oracle's order, bit for bit; and the device's
// PROX contact reduction of them must keep what the oracle's std::sort +
// std::unique keeps of the same contacts collected in callback order, as
// PROX collects them.  With --series FILE: PROX's loop over a motion series
// (run_series).  Exit status 0 when all match.  Build: python
// bvh-tvcg18_amd/build.py --oracle.
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kdop_amd.hpp"
#include "kdop_meshgen.h"

// oracle/liboracle.so (the CPU restatement of the sequential reference)
extern "C" {
struct OrcContact {
  uint32_t ta, tb;
  float p[3], n[3], d;
};
void* orc_tree_build(int K, uint64_t mem_bytes, uint32_t nt, const uint32_t* tets, const float* X, const float* Y, const float* Z);
void orc_tree_free(void* h);
void orc_tree_refit(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z);
void orc_transform_f32(uint32_t nv, const float* X0, const float* Y0, const float* Z0, const float* p, const float* q, float* X, float* Y,
                       float* Z);
int orc_traverse(void* hA, int ntA, const uint32_t* tetsA, const float* XA, const float* YA, const float* ZA, const uint8_t* surfA, void* hB,
                 int ntB, const uint32_t* tetsB, const float* XB, const float* YB, const float* ZB, const uint8_t* surfB, int do_narrow,
                 int threads, uint32_t* p1, uint64_t p1_cap, uint32_t* p2, uint64_t p2_cap, OrcContact* contacts, uint64_t c_cap,
                 uint64_t* counters, double* seconds);
uint64_t orc_prox_reduce(uint64_t n, const uint32_t* bi, const uint32_t* bj, const float* p, const float* nrm, const float* d, uint32_t* out);
int orc_raycast(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z, const uint8_t* surf, uint32_t n,
                const float* origins, const float* directions, float* hits, float* lengths, uint8_t* did_hit);
struct OrcSphereContact {
  uint32_t sphere, tet;
  float p[3], n[3], d;
};
uint64_t orc_single_traversal(void* h, const uint32_t* tets, const float* X, const float* Y, const float* Z, const uint8_t* surf, uint32_t n,
                              const float* spheres, int flip, OrcSphereContact* out, uint64_t cap);
}

struct MeshData {
  std::vector<float> X, Y, Z;
  std::vector<uint32_t> tets;
  std::vector<uint8_t> surf;
};

static MeshData box(uint32_t n, float size, double jitter, uint64_t seed) {
  uint32_t nv = 0, nt = 0;
  kdop_box_mesh_size(n, &nv, &nt);
  MeshData raw, m;
  raw.X.resize(nv); raw.Y.resize(nv); raw.Z.resize(nv); raw.tets.resize(4 * (size_t)nt);
  kdop_make_box_mesh(n, size, jitter, seed, raw.X.data(), raw.Y.data(), raw.Z.data(), raw.tets.data());
  m.X.resize(nv); m.Y.resize(nv); m.Z.resize(nv); m.tets.resize(4 * (size_t)nt); m.surf.resize(nt);
  kdop_mesh_reorder(nv, raw.X.data(), raw.Y.data(), raw.Z.data(), nt, raw.tets.data(), m.X.data(), m.Y.data(), m.Z.data(), m.tets.data());
  kdop_compute_surface_map(nv, nt, m.tets.data(), m.surf.data());
  return m;
}

struct Collect : kdop_amd::ContactsCallback {
  std::vector<float> v;  // 7 floats per contact: p, n, depth
  void operator()(const float p[3], const float n[3], float depth) override {
    v.insert(v.end(), p, p + 3);
    v.insert(v.end(), n, n + 3);
    v.push_back(depth);
  }
};

// PROX's collision-detection loop (prox_collision_detection.h:124-345) over a
// motion series written by tests/test_gpu_glue.py: per step the pool refit of
// every body (narrow::update_kdop_bvh, narrow_update_kdop_bvh.h:126-176), one
// tandem traversal of the pair list with a ContactsCallback per pair
// (dispatch_tetramesh_tetramesh, narrow_tetramesh_tetramesh.h:61-91), then the
// contact reduction of everything the callbacks collected.  Every step: each
// pair's contacts equal the oracle's (count, order, bits) and the reduction
// keeps what the oracle's std::sort + std::unique keeps.
// File: int32 K, nmeshes, npairs, steps; per mesh uint32 nv, nt, float X, Y, Z
// [nv], uint32 tets [4 nt], uint8 surface [nt]; uint32 pairs [2 npairs]; float
// P [steps][nmeshes][3], Q [steps][nmeshes][4].
static int run_series(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) { std::printf("cannot open %s\n", path); return 4; }
  auto rd = [&](void* dst, size_t bytes) {
    if (std::fread(dst, 1, bytes, f) != bytes) throw kdop_amd::Error(-1, "short series file");
  };
  int32_t hdr[4];
  int bad = 0;
  try {
    rd(hdr, sizeof hdr);
    const int K = hdr[0], nm = hdr[1], np = hdr[2], steps = hdr[3];
    std::vector<MeshData> meshes(nm);
    for (MeshData& m : meshes) {
      uint32_t nv = 0, nt = 0;
      rd(&nv, 4); rd(&nt, 4);
      m.X.resize(nv); m.Y.resize(nv); m.Z.resize(nv); m.tets.resize(4 * (size_t)nt); m.surf.resize(nt);
      rd(m.X.data(), 4 * (size_t)nv); rd(m.Y.data(), 4 * (size_t)nv); rd(m.Z.data(), 4 * (size_t)nv);
      rd(m.tets.data(), 16 * (size_t)nt); rd(m.surf.data(), nt);
    }
    std::vector<uint32_t> pairs(2 * (size_t)np);
    rd(pairs.data(), 8 * (size_t)np);
    std::vector<float> P(3 * (size_t)steps * nm), Q(4 * (size_t)steps * nm);
    rd(P.data(), 4 * P.size()); rd(Q.data(), 4 * Q.size());
    std::fclose(f);
    f = nullptr;
    kdop_amd::Context ctx(K);
    std::vector<kdop_amd::Tree> trees;
    std::vector<void*> orc;
    for (const MeshData& m : meshes) {
      trees.push_back(ctx.make_tree(m.X, m.Y, m.Z, m.tets, m.surf));
      orc.push_back(orc_tree_build(K, 8000, (uint32_t)m.surf.size(), m.tets.data(), m.X.data(), m.Y.data(), m.Z.data()));
    }
    size_t total = 0, kept_total = 0, live_steps = 0;
    std::vector<MeshData> world(nm);
    std::vector<OrcContact> want(1 << 14);
    for (int step = 0; step < steps; ++step) {
      std::vector<kdop_amd::UpdateWorkItem> pool(nm);
      for (int i = 0; i < nm; ++i) {
        pool[i].tree = &trees[i];
        std::memcpy(pool[i].p, &P[3 * ((size_t)step * nm + i)], 12);
        std::memcpy(pool[i].q, &Q[4 * ((size_t)step * nm + i)], 16);
        const MeshData& m = meshes[i];
        world[i].X.resize(m.X.size()); world[i].Y.resize(m.X.size()); world[i].Z.resize(m.X.size());
        orc_transform_f32((uint32_t)m.X.size(), m.X.data(), m.Y.data(), m.Z.data(), pool[i].p, pool[i].q, world[i].X.data(),
                          world[i].Y.data(), world[i].Z.data());
        orc_tree_refit(orc[i], m.tets.data(), world[i].X.data(), world[i].Y.data(), world[i].Z.data());
      }
      ctx.update_kdop_bvh(pool);
      std::vector<Collect> got(np);
      std::vector<kdop_amd::TestPair> tps;
      for (int k = 0; k < np; ++k) tps.push_back(kdop_amd::TestPair{&trees[pairs[2 * k]], &trees[pairs[2 * k + 1]], &got[k]});
      ctx.tandem_traversal(tps);
      const std::vector<kdop_prox_contact> red = ctx.reduce_contacts();
      // the oracle, pair by pair, and PROX's contacts vector in callback order
      size_t step_contacts = 0, pairs_ok = 0;
      std::vector<uint32_t> bi, bj;
      std::vector<float> Pc, Nc, Dc;
      for (int k = 0; k < np; ++k) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        uint64_t cnt[9];
        for (int pass = 0; pass < 2; ++pass) {  // again with room for all when the buffer was short
          orc_traverse(orc[a], 0, meshes[a].tets.data(), world[a].X.data(), world[a].Y.data(), world[a].Z.data(), meshes[a].surf.data(),
                       orc[b], 0, meshes[b].tets.data(), world[b].X.data(), world[b].Y.data(), world[b].Z.data(), meshes[b].surf.data(), 1, 1,
                       nullptr, 0, nullptr, 0, want.data(), want.size(), cnt, nullptr);
          if ((size_t)cnt[6] <= want.size()) break;
          want.resize((size_t)cnt[6]);
        }
        const size_t nw = (size_t)cnt[6], ng = got[k].v.size() / 7;
        bool same = nw == ng;
        for (size_t i = 0; same && i < nw; ++i) {
          const float w7[7] = {want[i].p[0], want[i].p[1], want[i].p[2], want[i].n[0], want[i].n[1], want[i].n[2], want[i].d};
          same = std::memcmp(w7, &got[k].v[7 * i], sizeof(w7)) == 0;
        }
        if (!same) std::printf("series step %d pair %d: %zu contacts (oracle %zu) MISMATCH\n", step, k, ng, nw);
        pairs_ok += same ? 1 : 0;
        step_contacts += ng;
        for (size_t i = 0; i < ng; ++i) {
          const float* v = &got[k].v[7 * i];
          bi.push_back(trees[a].body);
          bj.push_back(trees[b].body);
          Pc.insert(Pc.end(), v, v + 3);
          Nc.insert(Nc.end(), v + 3, v + 6);
          Dc.push_back(v[6]);
        }
      }
      std::vector<uint32_t> keep(bi.size() + 1);
      const size_t nk = (size_t)orc_prox_reduce(bi.size(), bi.data(), bj.data(), Pc.data(), Nc.data(), Dc.data(), keep.data());
      bool rsame = red.size() == nk;
      for (size_t i = 0; rsame && i < nk; ++i) {
        const uint32_t j = keep[i];
        const kdop_prox_contact& r = red[i];
        const float w7[7] = {Pc[3 * j], Pc[3 * j + 1], Pc[3 * j + 2], Nc[3 * j], Nc[3 * j + 1], Nc[3 * j + 2], Dc[j]};
        const float g7[7] = {r.p[0], r.p[1], r.p[2], r.n[0], r.n[1], r.n[2], r.depth};
        rsame = r.body_i == bi[j] && r.body_j == bj[j] && std::memcmp(w7, g7, sizeof(w7)) == 0;
      }
      const bool ok = pairs_ok == (size_t)np && rsame;
      std::printf("series step %d: %zu contacts, %zu kept, %zu of %d pairs and the reduction %s\n", step, step_contacts, red.size(),
                  pairs_ok, np, ok ? "match" : "MISMATCH");
      bad += ok ? 0 : 1;
      total += step_contacts;
      kept_total += red.size();
      live_steps += step_contacts ? 1 : 0;
    }
    for (void* h : orc) orc_tree_free(h);
    std::printf("series: %d steps, %zu with contacts, %zu contacts, %zu kept\n", steps, live_steps, total, kept_total);
    if (live_steps < (size_t)steps) {
      std::printf("steps without contacts: the series does not test every step\n");
      return 2;
    }
    return bad ? 1 : 0;
  } catch (const kdop_amd::Error& e) {
    if (f) std::fclose(f);
    std::printf("error %d: %s\n", e.code, e.what());
    return 3;
  }
}

int main(int argc, char** argv) {
  if (argc == 3 && std::strcmp(argv[1], "--series") == 0) return run_series(argv[2]);
  const int K = 26;
  std::vector<MeshData> meshes = {box(8, 1.0f, 0.1, 1), box(10, 1.2f, 0.1, 2), box(6, 0.8f, 0.1, 3)};
  const uint32_t pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
  std::vector<float> quats(4 * meshes.size() * 2);
  kdop_rng_quaternions(21, 0, (uint64_t)(meshes.size() * 2), quats.data());
  try {
    kdop_amd::Context ctx(K);
    std::vector<kdop_amd::Tree> trees;
    std::vector<void*> orc;
    for (const MeshData& m : meshes) {
      trees.push_back(ctx.make_tree(m.X, m.Y, m.Z, m.tets, m.surf));
      orc.push_back(orc_tree_build(K, 8000, (uint32_t)m.surf.size(), m.tets.data(), m.X.data(), m.Y.data(), m.Z.data()));
    }
    int bad = 0;
    size_t total = 0;
    for (int step = 0; step < 2; ++step) {
      std::vector<kdop_amd::UpdateWorkItem> pool;
      std::vector<MeshData> world(meshes.size());
      for (size_t i = 0; i < meshes.size(); ++i) {
        kdop_amd::UpdateWorkItem w;
        w.tree = &trees[i];
        // side by side with partial overlap (surfaces facing each other), moving a little per step
        static const float base[3][3] = {{0.0f, 0.0f, 0.0f}, {0.95f, 0.05f, 0.0f}, {0.3f, 0.75f, 0.1f}};
        w.p[0] = base[i][0] + 0.02f * (float)step; w.p[1] = base[i][1]; w.p[2] = base[i][2] - 0.01f * (float)step;
        std::memcpy(w.q, &quats[4 * (i + meshes.size() * step)], 4 * sizeof(float));
        pool.push_back(w);
        const MeshData& m = meshes[i];
        world[i].X.resize(m.X.size()); world[i].Y.resize(m.X.size()); world[i].Z.resize(m.X.size());
        orc_transform_f32((uint32_t)m.X.size(), m.X.data(), m.Y.data(), m.Z.data(), w.p, w.q, world[i].X.data(), world[i].Y.data(),
                          world[i].Z.data());
        orc_tree_refit(orc[i], m.tets.data(), world[i].X.data(), world[i].Y.data(), world[i].Z.data());
      }
      ctx.update_kdop_bvh(pool);
      std::vector<Collect> got(3);
      std::vector<kdop_amd::TestPair> tps;
      for (int k = 0; k < 3; ++k) tps.push_back(kdop_amd::TestPair{&trees[pairs[k][0]], &trees[pairs[k][1]], &got[k]});
      ctx.tandem_traversal(tps);
      for (int k = 0; k < 3; ++k) {
        const uint32_t a = pairs[k][0], b = pairs[k][1];
        std::vector<OrcContact> want(1 << 16);
        uint64_t cnt[9];
        orc_traverse(orc[a], 0, meshes[a].tets.data(), world[a].X.data(), world[a].Y.data(), world[a].Z.data(), meshes[a].surf.data(), orc[b], 0,
                     meshes[b].tets.data(), world[b].X.data(), world[b].Y.data(), world[b].Z.data(), meshes[b].surf.data(), 1, 1, nullptr, 0,
                     nullptr, 0, want.data(), want.size(), cnt, nullptr);
        const size_t nw = (size_t)cnt[6];
        const size_t ng = got[k].v.size() / 7;
        bool same = nw == ng && nw <= want.size();
        for (size_t i = 0; same && i < nw; ++i) {
          const float w7[7] = {want[i].p[0], want[i].p[1], want[i].p[2], want[i].n[0], want[i].n[1], want[i].n[2], want[i].d};
          same = std::memcmp(w7, &got[k].v[7 * i], sizeof(w7)) == 0;
        }
        std::printf("step %d pair %d: %zu contacts (oracle %zu) %s\n", step, k, ng, nw, same ? "match" : "MISMATCH");
        bad += same ? 0 : 1;
        total += ng;
      }
      // PROX's contacts vector: every pair's callback contacts, pair by pair
      std::vector<uint32_t> bi, bj;
      std::vector<float> P, N, D;
      for (int k = 0; k < 3; ++k)
        for (size_t i = 0; i < got[k].v.size() / 7; ++i) {
          const float* v = &got[k].v[7 * i];
          bi.push_back(trees[pairs[k][0]].body);
          bj.push_back(trees[pairs[k][1]].body);
          P.insert(P.end(), v, v + 3);
          N.insert(N.end(), v + 3, v + 6);
          D.push_back(v[6]);
        }
      std::vector<uint32_t> keep(bi.size() + 1);
      const size_t nk = (size_t)orc_prox_reduce(bi.size(), bi.data(), bj.data(), P.data(), N.data(), D.data(), keep.data());
      const std::vector<kdop_prox_contact> red = ctx.reduce_contacts();
      bool same = red.size() == nk;
      for (size_t i = 0; same && i < nk; ++i) {
        const uint32_t j = keep[i];
        const kdop_prox_contact& r = red[i];
        const float w7[7] = {P[3 * j], P[3 * j + 1], P[3 * j + 2], N[3 * j], N[3 * j + 1], N[3 * j + 2], D[j]};
        const float g7[7] = {r.p[0], r.p[1], r.p[2], r.n[0], r.n[1], r.n[2], r.depth};
        same = r.body_i == bi[j] && r.body_j == bj[j] && std::memcmp(w7, g7, sizeof(w7)) == 0;
      }
      std::printf("step %d reduction: %zu of %zu contacts kept (oracle %zu) %s\n", step, red.size(), bi.size(), nk, same ? "match" : "MISMATCH");
      bad += same ? 0 : 1;
      if (step == 1) {
        // kdop::raycast and kdop::single_traversal on body 0 in this pose: one query at a time and in one batch
        const MeshData& m0 = meshes[0];
        const uint32_t nr = 64;
        std::vector<float> o(3 * nr), d(3 * nr), sp(4 * nr);
        for (uint32_t i = 0; i < nr; ++i) {
          const float a = 0.37f * (float)i, z = -0.9f + 1.8f * (float)i / (float)nr;
          const float c = std::sqrt(1.0f - z * z);
          const float u[3] = {c * std::cos(a), c * std::sin(a), z};
          for (int k = 0; k < 3; ++k) {
            o[3 * i + k] = pool[0].p[k] + 2.0f * u[k];
            d[3 * i + k] = -u[k] + 0.05f * (float)((i + k) % 5 - 2);
            sp[4 * i + k] = pool[0].p[k] + 0.6f * u[k];
          }
          sp[4 * i + 3] = 0.02f + 0.002f * (float)(i % 50);
        }
        std::vector<float> hw(3 * nr), lw(nr), hg(3 * nr), lg(nr);
        std::vector<uint8_t> fw(nr), fg(nr);
        orc_raycast(orc[0], m0.tets.data(), world[0].X.data(), world[0].Y.data(), world[0].Z.data(), m0.surf.data(), nr, o.data(), d.data(),
                    hw.data(), lw.data(), fw.data());
        ctx.raycast(trees[0], nr, o.data(), d.data(), hg.data(), lg.data(), fg.data());
        bool rsame = std::memcmp(hw.data(), hg.data(), hw.size() * 4) == 0 && std::memcmp(lw.data(), lg.data(), lw.size() * 4) == 0 &&
                     fw == fg;
        uint32_t nhit = 0;
        for (uint32_t i = 0; i < nr && rsame; ++i) {
          float h1[3], l1 = 0.0f;
          const bool hit1 = ctx.raycast(trees[0], &o[3 * i], &d[3 * i], h1, l1);
          rsame = hit1 == (fw[i] != 0) && std::memcmp(h1, &hw[3 * i], 12) == 0 && std::memcmp(&l1, &lw[i], 4) == 0;
          nhit += hit1 ? 1u : 0u;
        }
        std::printf("raycast: %u of %u rays hit %s\n", nhit, nr, rsame ? "match" : "MISMATCH");
        bad += rsame ? 0 : 1;
        std::vector<OrcSphereContact> sw(1 << 16);
        const size_t ns = (size_t)orc_single_traversal(orc[0], m0.tets.data(), world[0].X.data(), world[0].Y.data(), world[0].Z.data(),
                                                       m0.surf.data(), nr, sp.data(), 0, sw.data(), sw.size());
        const std::vector<kdop_sphere_contact> sg = ctx.single_traversal(trees[0], nr, sp.data());
        bool ssame = ns == sg.size() && ns <= sw.size();
        for (size_t i = 0; ssame && i < ns; ++i)
          ssame = sw[i].sphere == sg[i].sphere && sw[i].tet == sg[i].tet && std::memcmp(sw[i].p, sg[i].p, 28) == 0;
        size_t k = 0;  // the same contacts through a callback, one sphere at a time
        for (uint32_t i = 0; i < nr && ssame; ++i) {
          Collect cb;
          ctx.single_traversal(&sp[4 * i], sp[4 * i + 3], trees[0], cb);
          for (size_t j = 0; ssame && j < cb.v.size() / 7; ++j, ++k)
            ssame = k < ns && sw[k].sphere == i && std::memcmp(sw[k].p, &cb.v[7 * j], 28) == 0;
        }
        ssame = ssame && k == ns;
        std::printf("single traversal: %zu contacts %s\n", ns, ssame ? "match" : "MISMATCH");
        bad += ssame ? 0 : 1;
      }
    }
    for (void* h : orc) orc_tree_free(h);
    if (total == 0) {
      std::printf("no contacts: the workload does not test anything\n");
      return 2;
    }
    return bad ? 1 : 0;
  } catch (const kdop_amd::Error& e) {
    std::printf("error %d: %s\n", e.code, e.what());
    return 3;
  }
}
