// kdop_hip.hip -- host side of the MI355X kDOP library: context, device
// resident bodies and trees, the refit / traversal / narrow-phase drivers and
// the C ABI declared in include/kdop_amd.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kdop_amd.h"
#include "kdop_kernels.h"
#include "kdop_persistent.h"
#include "kdop_adjoint.h"
#include "kdop_sort.h"
#include "kdop_raycast.h"
#include "kdop_single.h"
#include "kdop_meshgen.h"
#include "tree_shape.h"

using namespace kdop_amd;

// Stack keep that never hands work on: above any stack, and small enough that
// keep x waves per workgroup cannot overflow.
constexpr uint32_t kNoHandOffKeep = 1u << 20;
// Default iteration budget of a persistent round: above the longest chain of
// a block on the 320k-tet pairs (38 iterations at K = 26, 52 at K = 14), so
// such a query runs one round, while a batch whose heavy pairs would keep a
// few blocks busy for hundreds of iterations is spread over later rounds.
constexpr uint32_t kRoundItersDefault = 64;
// ... and for calls of at least kRoundItersBatchPairs pairs, whose rounds
// hand heavy stacks on anyway: 512-pair batch at 64 / 48 / 32 / 24 / 16
// iterations 1.544-1.555 / 1.555-1.559 / 1.522-1.532 / 1.507-1.509 /
// 1.518-1.519 ms per call, 64-pair refit series step 0.859 / 0.854-0.857 /
// 0.856-0.857 ms at 64 / 32 / 24 (a one-pair query at 24 hands its tail on:
// 0.186 -> 0.206 ms per step).
constexpr uint32_t kRoundItersBatch = 24;
constexpr uint32_t kRoundItersBatchPairs = 64;
// Calls of at least kManyPairs pairs (many small trees) take a wider ramp
// target and a larger seed budget by default: 512-pair batch with 65,536 /
// 262,144 / 1,048,576 seeds 1.457-1.461 / 1.415-1.419 / 1.444-1.446 ms per
// call, a 64-pair refit series step 0.856-0.857 against 0.931-0.933 ms and a
// one-pair query 0.189 against 0.205 ms with 262,144.
constexpr uint32_t kManyPairs = 256;
constexpr uint32_t kSeedBudgetDefault = 65536, kSeedBudgetManyPairs = 262144;

namespace {

struct BodyHost {
  uint32_t nv = 0, nt = 0, L = 0, H = 0, depth = 0;
  std::vector<uint32_t> C, E, level_base;  // per level (0 = top)
  uint32_t nnodes = 0, node_base = 0;  // slots (nodes + free slots), first slot
  uint64_t real_nodes = 0;             // tree nodes below the root record
  uint32_t top_base = 0, ntop = 0;
  std::vector<uint32_t> roots_off;  // per level: offset of the level's chunk roots in the context's root table
  float *X0 = nullptr, *Y0 = nullptr, *Z0 = nullptr, *X = nullptr, *Y = nullptr, *Z = nullptr;
  float4* W = nullptr;  // world vertices as (x, y, z, 0)
  float4* W0 = nullptr;  // rest positions as (x, y, z, 0)
  uint32_t* tets = nullptr;
  uint8_t* surf = nullptr;
  // chunks wider than 32 leaves (topology built on the host): each tet's leaf
  // slot, and the internal nodes ordered by height (union_off[h] .. union_off[h
  // + 1]: the nodes of height h + 1)
  uint32_t* leaf_slot = nullptr;
  uint32_t* unions = nullptr;
  std::vector<uint32_t> union_off;
  bool wide() const { return L > 32; }
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
};

// The launches of one refit batch, derived from its body list alone (the
// plan's offsets in the device copy): kept on the context and reused while
// the list repeats, so a repeated refit makes its first launch at once.
struct RefitStep {
  size_t o_bb, o_gpre, o_gitem, o_cfirst, o_ncnk, o_cbase, o_cn;
  uint32_t m, blocks;
};
struct RefitLaunch {
  bool valid = false;
  size_t o_bb = 0, o_vpre = 0, o_tpre = 0, o_vwave = 0, o_twave = 0, o_pbody = 0, o_pchunk = 0, o_top = 0;
  uint64_t total_v = 0, total_leaf_threads = 0;
  uint32_t npbody = 0;
  bool blk8 = false;  // every body has L = 8: 512-thread leaves blocks, which refit level H-3 too
  std::vector<RefitStep> steps;
  std::vector<uint32_t> wide;  // bodies refitted from their own lists (chunks over 32 leaves)
};

}  // namespace

struct kdop_ctx {
  kdop_config cfg{};
  int K = 8;
  DirTable dt{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  std::vector<BodyHost> bodies;
  // node pool (records of K+2 floats), body descriptors, top-level roots
  DevBuf<float> nodes;
  size_t nodes_used = 0;
  DevBuf<BodyDev> dbodies;
  DevBuf<float> dposes;  // 8 floats per body: the pose of its last refit (p, q xyzw, 0)
  std::vector<uint32_t> tops;
  DevBuf<uint32_t> dtops;
  // traversal state
  DevBuf<uint2> dpairs;
  DevBuf<SeedDesc> dseeds;
  uint32_t seed_budget = kSeedBudgetDefault;  // seeds (chunk-root pairs tested first) per call, when set
  bool seed_budget_set = false;                 // the caller set it; else by the call's pair count
  uint32_t seed_budget_eff = kSeedBudgetDefault;  // the budget of the current (or last) call
  DevBuf<Item> front[2];
  DevBuf<LeafPair> p1, p2;
  DevBuf<kdop_contact> contacts;
  DevBuf<kdop_contact> contacts_tmp;  // kdop_sort_contacts: the permuted copy
  DevBuf<unsigned char> sortbuf;       // kdop_sort_contacts: two key arrays
  DevBuf<unsigned char> raybuf;        // kdop_raycast / kdop_single_traversal: inputs and counts
  DevBuf<kdop_sphere_contact> sphere_out;  // kdop_single_traversal: the contacts
  DevBuf<uint2> sg_items[2];               // kdop_single_traversal: (sphere, node) frontiers
  DevBuf<uint32_t> sg_scan;                // kdop_single_traversal / kdop_raycast: counts, offsets, block sums, total, flag
  DevBuf<uint8_t> sg_kind;                 // kdop_single_traversal: what each item of a two-level step becomes
  DevBuf<float> rc_best;                   // kdop_raycast: each (ray, chunk) item's best length from FLT_MAX
  DevBuf<uint32_t> rc_first;               // kdop_raycast: each ray's first item, then a flag
  uint32_t sg_max_items = 1u << 28;        // KDOP_OPT_SINGLE_ITEMS
  DevBuf<uint32_t> redbuf;                 // contact reduction: keep flags, offsets, scan sums, total
  DevBuf<kdop_prox_contact> prox_io;       // kdop_reduce_prox_contacts: staged input and output
  uint32_t reduced_n = 0;                  // contacts left by kdop_reduce_contacts
  bool reduced = false;                    // kdop_reduce_contacts ran since the last traversal
  DevBuf<NarrowRec> recs;
  DevBuf<uint32_t> gate;       // two-pass narrow phase: surface pairs through the SAT gate
  uint64_t last_p2 = 0;        // surface pairs of the last traversal (picks the narrow path)
  int narrow_path = 0;         // KDOP_OPT_NARROW_PATH
  bool use_graphs = false;     // KDOP_OPT_GRAPHS: replay the traversal's launches from a HIP graph
  std::vector<uint64_t> tg_sig;    // signature (every launch argument) of the captured traversal
  std::vector<uint64_t> prev_sig;  // signature of the previous call (capture once it repeats)
  hipGraph_t tg = nullptr;
  hipGraphExec_t tg_exec = nullptr;
  DevBuf<double> adj_acc;      // adjoint accumulators [npairs][kAdjStripes][kAdjVals]
  DevBuf<uint32_t> adj_bodies;
  CallState* dstate = nullptr;    // counters and frontier segments of one call
  CallSummary* hsum = nullptr;    // call summary, pinned host memory written by k_sum_counters
  CallSummary* hsum_dev = nullptr;  // its device address
  HostWords* hwords = nullptr;      // step sizes of the single traversal / raycast, pinned host memory
  HostWords* hwords_dev = nullptr;  // its device address
  uint32_t hwords_seq = 0;
  uint32_t last_npairs = 0;
  bool have_traversal = false, have_p1 = false, have_narrow = false;
  std::vector<uint2> last_pairs;
  // refit scratch
  DevBuf<uint32_t> plan;
  DevBuf<float> poses;
  int num_cus = 0;
  // persistent launch shape as set (0: the default of the kernel that runs; see ls_of / wpc_of / bw_of)
  int ls = 0, wpc = 0;
  int bw = 0;                // waves per persistent workgroup (1: a stack per wave; 2, 4, 8: a shared stack)
  int item_lanes = 4;        // lanes per work item in the persistent rounds (4: k_persist_quad; 1: k_persist / k_persist_blk)
  int quad_items = 1;        // KDOP_OPT_QUAD_ITEMS: items per quad per iteration of k_persist_quad (1 or 2)
  int node_slab = 0;         // KDOP_OPT_NODE_SLAB: 1 stages each wave's records in an LDS node slab (k_persist_quad SLAB)
  uint32_t round_iters = kRoundItersDefault;  // KDOP_OPT_ROUND_ITERS: a block's iteration budget per round once the frontier is gone (0: none)
  bool round_iters_set = false;               // the caller set it; else by the call's pair count
  uint32_t round_iters_eff = kRoundItersDefault;  // the budget of the current (or last) call
  uint32_t ramp_target = 0;  // frontier width that ends the breadth-first levels (0: 16 x persistent waves)
  uint32_t bfs_max = 0;      // most breadth-first levels before the persistent rounds (0: by the call's size, bfs_auto)
  // the level at which the last persistent call's ramp stopped (an empty or
  // target-wide frontier; ~0u: it ran to its last level), the seeds and
  // target it ran with: a call from the same seeds queues no levels past it
  uint64_t seed_gen = 0;  // bumped whenever the seed descriptors are rebuilt
  uint64_t ramp_seed_gen = ~0ull;
  uint32_t ramp_stop = 0, ramp_stop_target = 0;
  int timing_level = 0;        // 0 none, 1 phase events, 2 + per-round dispatch events
  uint32_t rounds_hint = 1;    // persistent rounds to queue next call (1..3), from the rounds recent calls needed
  uint32_t rounds_calm = 0;    // consecutive calls that needed fewer rounds than rounds_hint
  uint32_t rounds_queued = 3;  // persistent rounds queued by the current call
  uint32_t max_rounds = kMaxRounds - 1;  // KDOP_OPT_MAX_ROUNDS: persistent rounds before the levels finish a spill
  bool async_refit = false;    // refit returns once queued (the next call orders after it)
  bool state_clean = false;    // CallState zeroed (by the last call's final fold)
  std::vector<uint32_t> plan_key;   // bodies of the refit plan in c->plan
  RefitLaunch rl;                   // its launches
  std::vector<uint2> seed_key;      // pair list whose pairs / seeds are on the device
  uint32_t seed_key_budget = 0;
  size_t seed_key_bodies = 0;
  uint64_t seed_total = 0;
  std::vector<uint64_t> seed_plan_est;  // per pair: seeds tested
  uint32_t keep = kNoHandOffKeep;  // items a wave keeps once the frontier is gone (default: no hand-off)
  uint32_t last_rounds = 0;
  bool last_persistent = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t evr[2 * kMaxRounds] = {};  // start / end of each persistent round
  double timing[4 + kMaxRounds] = {};  // see kdop_last_timing
  uint32_t last_levels = 0;
  uint32_t last_fallback_levels = 0;  // levels launched after the persistent rounds ran out (last call)
  bool count_records = false;         // KDOP_OPT_COUNT_RECORDS: mark the records the persistent rounds read
  DevBuf<uint8_t> touched;            // ... one byte per node slot
  uint64_t last_distinct_records = 0; // distinct records the last counting call's rounds read
  void* scratch = nullptr;            // this device's address of the allocation-free entry point's array
  std::string scratch_err;            // why it is null
};

namespace {

// Persistent launch shape in effect: the option's value, or the default of the
// kernel the item-lane setting selects (quad kernel: 4-wave workgroups, 24
// waves per CU, 256 stack entries per wave; one lane per item: 2-wave
// workgroups, 8 waves per CU, 512 entries per wave).
// quad rounds: 512-item stacks shared by 4 waves (16.4 KB of LDS, 70 VGPRs at K = 26), 7 workgroups per CU
// (C2: round 0 68.8 -> 65.3 us against 1,024-item stacks at 6 per CU)
int ls_of(const kdop_ctx* c) { return c->ls ? c->ls : (c->item_lanes == 4 ? 128 : 512); }
int wpc_of(const kdop_ctx* c) { return c->wpc ? c->wpc : (c->item_lanes == 4 ? 28 : 8); }
int bw_of(const kdop_ctx* c) { return c->bw ? c->bw : (c->item_lanes == 4 ? 4 : 2); }

int fail(kdop_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                               \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return fail((ctx), KDOP_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                             \
  } while (0)

template <class T>
int ensure(kdop_ctx* c, DevBuf<T>& b, size_t n, bool keep = false) {
  if (b.cap >= n) return KDOP_OK;
  size_t cap = std::max<size_t>(n, b.cap * 2);
  T* p = nullptr;
  if (hipMalloc(&p, cap * sizeof(T)) != hipSuccess) {
    cap = n;
    hipError_t e = hipMalloc(&p, cap * sizeof(T));
    if (e != hipSuccess) return fail(c, KDOP_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", cap * sizeof(T), hipGetErrorString(e));
  }
  if (keep && b.p && b.cap) {
    HIPCHK(c, hipMemcpyAsync(p, b.p, b.cap * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (b.p) (void)hipFree(b.p);
  b.p = p;
  b.cap = cap;
  return KDOP_OK;
}
template <class T>
void release(DevBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

bool valid_K(int K) { return K == 6 || K == 8 || K == 12 || K == 14 || K == 18 || K == 20 || K == 26; }

// geometry_direction_table.h:53-172, computed on the host in float so the
// device sees the reference's bits.
void make_dir_table(int K, DirTable& dt) {
  auto unit = [](float x, float y, float z, float* o) {
    float d = 0.0f;
    d = d + x * x;
    d = d + y * y;
    d = d + z * z;
    const float l = std::sqrt(d);
    const float inv = 1.0f / l;
    o[0] = x * inv; o[1] = y * inv; o[2] = z * inv;
  };
  float d3[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  float d4[4][3], d6[6][3];
  const float raw4[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
  const float raw6[6][3] = {{1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {0, 1, 1}, {-1, 0, 1}, {0, -1, 1}};
  for (int i = 0; i < 4; ++i) unit(raw4[i][0], raw4[i][1], raw4[i][2], d4[i]);
  for (int i = 0; i < 6; ++i) unit(raw6[i][0], raw6[i][1], raw6[i][2], d6[i]);
  int n = 0;
  auto put = [&](float (*src)[3], int cnt) {
    for (int i = 0; i < cnt; ++i, ++n) for (int j = 0; j < 3; ++j) dt.d[3 * n + j] = src[i][j];
  };
  switch (K / 2) {
    case 3: put(d3, 3); break;
    case 4: put(d4, 4); break;
    case 6: put(d6, 6); break;
    case 7: put(d3, 3); put(d4, 4); break;
    case 9: put(d3, 3); put(d6, 6); break;
    case 10: put(d4, 4); put(d6, 6); break;
    case 13: put(d3, 3); put(d4, 4); put(d6, 6); break;
  }
}

// ----------------------------------------------------------- K dispatch
template <template <int> class F, class... Args>
void dispatch_K(int K, Args&&... args) {
  switch (K) {
    case 6: F<6>::run(args...); break;
    case 8: F<8>::run(args...); break;
    case 12: F<12>::run(args...); break;
    case 14: F<14>::run(args...); break;
    case 18: F<18>::run(args...); break;
    case 20: F<20>::run(args...); break;
    case 26: F<26>::run(args...); break;
  }
}

uint32_t host_chunk_elems(const BodyHost& b, int h, uint32_t c) { return (c + 1 < b.C[h]) ? b.L : b.E[h] - c * b.L; }
uint32_t host_stride(const BodyHost& b, int h) { return (h == (int)b.H - 1) ? 2 * b.L : b.L; }
uint32_t host_resolved_root(const BodyHost& b, int h, uint32_t c) {
  for (;;) {
    const uint32_t n = host_chunk_elems(b, h, c);
    if (h == (int)b.H - 1 || n >= 2) return b.node_base + b.level_base[h] + c * host_stride(b, h) + 1;
    c = c * b.L;
    ++h;
  }
}

// The median-split topology of a body whose chunks hold more than 32 leaves
// (kdop_make_tree.h:57-121 branch chunks, :123-186 super chunks), in the
// slot layout the device build writes for narrower chunks: per chunk, BFS
// over element ranges; branch chunks store every range, super chunks only
// ranges of two or more elements (a single element resolves to the child
// chunk's root).  Fills meta (2 words per slot of the body), each tet's leaf
// slot, and the internal nodes grouped by height for the refit.
void build_wide_topology(const BodyHost& b, uint32_t body, const uint8_t* surf, std::vector<uint32_t>& meta,
                         std::vector<uint32_t>& leaf_slot, std::vector<uint32_t>& unions, std::vector<uint32_t>& union_off) {
  meta.assign(2 * (size_t)b.nnodes, 0xFFFFFFFFu);
  meta[0] = meta[1] = 0u;  // the root record
  leaf_slot.assign(b.nt, 0u);
  std::vector<uint32_t> qlo, qhi;
  for (uint32_t h = 0; h < b.H; ++h) {
    const bool branch = h == b.H - 1;
    for (uint32_t ch = 0; ch < b.C[h]; ++ch) {
      const uint32_t n = host_chunk_elems(b, (int)h, ch);
      if (!branch && n < 2) continue;
      const uint32_t base = b.level_base[h] + ch * host_stride(b, (int)h) + 1;  // body-local slot of BFS node 0
      const uint32_t first = ch * b.L;
      qlo.assign(1, 0u);
      qhi.assign(1, n - 1u);
      for (size_t head = 0; head < qlo.size(); ++head) {
        const uint32_t lo = qlo[head], hi = qhi[head], local = base + (uint32_t)head;
        const uint32_t rootflag = (branch && head == 0) ? F_ROOT : 0u;
        if (lo == hi) {  // branch leaf
          const uint32_t tet = first + lo;
          meta[2 * (size_t)local] = F_LEAF | F_BRANCH | tet;
          meta[2 * (size_t)local + 1] = rootflag | ((surf[tet] & 15u) ? F_SURF : 0u) | body;
          leaf_slot[tet] = b.node_base + local;
          continue;
        }
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t rlo[2] = {lo, mid + 1u}, rhi[2] = {mid, hi};
        uint32_t child[2];
        for (int s2 = 0; s2 < 2; ++s2) {
          if (branch || rlo[s2] < rhi[s2]) {
            child[s2] = b.node_base + base + (uint32_t)qlo.size();
            qlo.push_back(rlo[s2]);
            qhi.push_back(rhi[s2]);
          } else {
            child[s2] = host_resolved_root(b, (int)h + 1, first + rlo[s2]);
          }
        }
        meta[2 * (size_t)local] = (branch ? F_BRANCH : 0u) | child[0];
        meta[2 * (size_t)local + 1] = rootflag | child[1];
      }
    }
  }
  // heights: children sit in later slots (within a chunk in BFS order, across
  // chunks on lower levels), so a reverse sweep sees them first
  std::vector<uint8_t> height(b.nnodes, 0);
  uint32_t hmax = 0;
  std::vector<uint32_t> count(1, 0);
  for (int64_t i = (int64_t)b.nnodes - 1; i >= 1; --i) {
    const uint32_t m0 = meta[2 * (size_t)i], m1 = meta[2 * (size_t)i + 1];
    if (m0 == 0xFFFFFFFFu || (m0 & F_LEAF)) continue;
    const uint32_t c0 = (m0 & IDX_MASK) - b.node_base, c1 = (m1 & IDX_MASK) - b.node_base;
    const uint32_t ht = 1u + std::max(height[c0], height[c1]);
    height[i] = (uint8_t)ht;
    if (ht > hmax) { hmax = ht; count.resize(hmax + 1, 0); }
    ++count[ht];
  }
  union_off.assign(hmax + 1, 0);
  for (uint32_t h = 1; h <= hmax; ++h) union_off[h] = union_off[h - 1] + count[h];
  unions.assign(union_off[hmax], 0u);
  std::vector<uint32_t> fill(union_off.begin(), union_off.end() - 1);
  for (uint32_t i = 1; i < b.nnodes; ++i) {
    const uint32_t m0 = meta[2 * (size_t)i];
    if (m0 == 0xFFFFFFFFu || (m0 & F_LEAF)) continue;
    unions[fill[height[i] - 1]++] = b.node_base + i;
  }
}

int upload_bodies(kdop_ctx* c) {
  std::vector<BodyDev> h(c->bodies.size());
  for (size_t i = 0; i < c->bodies.size(); ++i) {
    const BodyHost& b = c->bodies[i];
    BodyDev& d = h[i];
    std::memset(&d, 0, sizeof(d));
    d.node_base = b.node_base; d.nnodes = b.nnodes; d.nt = b.nt; d.nv = b.nv;
    d.top_base = b.top_base; d.ntop = b.ntop; d.nbranch = b.C[b.H - 1]; d.H = b.H; d.L = b.L;
    d.branch_base = b.node_base + b.level_base[b.H - 1];
    d.super_base = b.H >= 2 ? b.node_base + b.level_base[b.H - 2] : 0u;
    d.super3_base = b.H >= 3 ? b.node_base + b.level_base[b.H - 3] : 0u;
    d.depth = b.depth;
    d.tets = b.tets; d.surf = b.surf; d.X0 = b.X0; d.Y0 = b.Y0; d.Z0 = b.Z0; d.W0 = b.W0; d.X = b.X; d.Y = b.Y; d.Z = b.Z; d.W = b.W;
  }
  int rc = ensure(c, c->dbodies, h.size());
  if (!rc) rc = ensure(c, c->dposes, 8 * h.size(), true);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->dbodies.p, h.data(), h.size() * sizeof(BodyDev), hipMemcpyHostToDevice, c->stream));
  rc = ensure(c, c->dtops, c->tops.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->dtops.p, c->tops.data(), c->tops.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

// ------------------------------------------------------------------- refit
template <int K>
struct RefitLeaves {
  static void run(kdop_ctx* c, uint32_t nthreads, bool blk8, const uint32_t* bb, const uint32_t* tpre, const uint32_t* vpre,
                  const uint32_t* twave, const uint32_t* vwave, uint32_t n, const float* pos, const float* quat, bool identity) {
    if (blk8)
      hipLaunchKernelGGL((k_refit_leaves<K, 8>), dim3((nthreads + 511) / 512), dim3(512), 0, c->stream, c->dbodies.p, bb, tpre, vpre, twave,
                         vwave, n, pos, quat, identity ? 1 : 0, c->dposes.p, c->nodes.p, c->dt);
    else
      hipLaunchKernelGGL((k_refit_leaves<K, 4>), dim3((nthreads + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, bb, tpre, vpre, twave,
                         vwave, n, pos, quat, identity ? 1 : 0, c->dposes.p, c->nodes.p, c->dt);
  }
};
// A body with chunks wider than 32 leaves: leaf DOPs at their slots, then
// its internal nodes height by height.
template <int K>
struct RefitWide {
  static void run(kdop_ctx* c, uint32_t body, const BodyHost& b) {
    hipLaunchKernelGGL(k_refit_tet_slots<K>, dim3((b.nt + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body,
                       (const uint32_t*)b.leaf_slot, c->nodes.p, c->dt);
    for (size_t h = 1; h < b.union_off.size(); ++h) {
      const uint32_t lo = b.union_off[h - 1], cnt = b.union_off[h] - lo;
      if (cnt)
        hipLaunchKernelGGL(k_refit_union_list<K>, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, (const uint32_t*)b.unions + lo, cnt,
                           c->nodes.p);
    }
  }
};
template <int K>
struct GatherProbe {
  static void run(kdop_ctx* c, const uint32_t* idx, uint32_t nidx, uint32_t* out, uint32_t blocks) {
    hipLaunchKernelGGL(k_gather_probe<K>, dim3(blocks), dim3(256), 0, c->stream, c->nodes.p, idx, nidx, out);
  }
};
template <int K>
struct RefitBranch {
  static void run(kdop_ctx* c, uint32_t nlist, const uint32_t* lbody, const uint32_t* lchunk) {
    hipLaunchKernelGGL(k_refit_branch<K>, dim3(nlist), dim3(64), 0, c->stream, c->dbodies.p, lbody, lchunk, nlist, c->nodes.p, c->dt);
  }
};
template <int K>
struct RefitSuper {
  static void run(kdop_ctx* c, uint32_t nblocks, const uint32_t* bb, const uint32_t* gpre, const uint32_t* gitem, const uint32_t* cfirst,
                  const uint32_t* ncnk, const uint32_t* cbase, const uint32_t* cn, uint32_t n) {
    hipLaunchKernelGGL(k_refit_super<K>, dim3(nblocks), dim3(64), 0, c->stream, c->dbodies.p, bb, gpre, gitem, cfirst, ncnk, cbase, cn, n,
                       c->nodes.p);
  }
};
template <int K>
struct RcSeedCount {
  static void run(kdop_ctx* c, uint32_t body, uint32_t n, const float* o, const float* d, uint32_t* count) {
    hipLaunchKernelGGL(k_rc_seed_count<K>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, c->nodes.p, c->dt, n, o, d,
                       count);
  }
};
template <int K>
struct RcExpand2Count {
  static void run(kdop_ctx* c, const uint2* items, uint32_t m, const float* o, const float* d, uint32_t* count, uint8_t* kind) {
    hipLaunchKernelGGL(k_rc_expand2_count<K>, dim3((m + 255) / 256), dim3(256), 0, c->stream, items, m, c->nodes.p, c->dt, o, d, count,
                       kind);
  }
};
template <int K>
struct RcRoots {
  static void run(kdop_ctx* c, const uint2* items, uint32_t m, const float* o, const float* d, uint32_t* pass, float* best, uint32_t* bad) {
    hipLaunchKernelGGL(k_rc_roots<K>, dim3((m + 255) / 256), dim3(256), 0, c->stream, items, m, c->nodes.p, c->dt, o, d, pass, best, bad);
  }
};
template <int K>
struct RcChunks {
  static void run(kdop_ctx* c, uint32_t body, const uint2* items, const uint32_t* list, uint32_t np, const float* o, const float* d,
                  float* best) {
    hipLaunchKernelGGL(k_rc_chunks<K>, dim3((np + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, items, list, np, c->nodes.p,
                       c->dt, o, d, best);
  }
};
template <int K>
struct RcResolve {
  static void run(kdop_ctx* c, uint32_t body, const uint2* items, uint32_t m, const float* best, const uint32_t* first, uint32_t n,
                  const float* o, const float* d, uint32_t ray_base, float* hits, float* lengths, uint8_t* flags) {
    hipLaunchKernelGGL(k_rc_resolve<K>, dim3((n + 3) / 4), dim3(256), 0, c->stream, c->dbodies.p, body, items, m, best, first, n,
                       c->nodes.p, c->dt, o, d, ray_base, hits, lengths, flags);
  }
};
template <int K>
struct SgSeedCount {
  static void run(kdop_ctx* c, uint32_t body, uint32_t n, const float* spheres, uint32_t* count) {
    hipLaunchKernelGGL(k_sg_seed_count<K>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, c->nodes.p, c->dt, n,
                       spheres, count);
  }
};
template <int K>
struct SgExpand2Count {
  static void run(kdop_ctx* c, const uint2* items, uint32_t m, const float* spheres, uint32_t* count, uint8_t* kind) {
    hipLaunchKernelGGL(k_sg_expand2_count<K>, dim3((m + 255) / 256), dim3(256), 0, c->stream, items, m, c->nodes.p, c->dt, spheres, count,
                       kind);
  }
};
template <int K>
struct SgExpand2Write {
  static void run(kdop_ctx* c, const uint2* items, uint32_t m, const uint8_t* kind, const uint32_t* off, uint2* out) {
    hipLaunchKernelGGL(k_sg_expand2_write<K>, dim3((m + 255) / 256), dim3(256), 0, c->stream, items, m, c->nodes.p, kind, off, out);
  }
};
template <int K>
struct SgLeaves {
  static void run(kdop_ctx* c, uint32_t body, const uint2* items, uint32_t m, const float* spheres, uint32_t sphere_base, int flip,
                  int write, uint32_t* flag, const uint32_t* off, kdop_sphere_contact* out, uint32_t* bad) {
    hipLaunchKernelGGL(k_sg_leaves<K>, dim3((m + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, items, m, c->nodes.p, c->dt,
                       spheres, sphere_base, flip, write, flag, off, out, bad);
  }
};
template <int K>
struct RefitTop {
  static void run(kdop_ctx* c, uint32_t n, const uint4* ent, const uint32_t* bb) {
    hipLaunchKernelGGL(k_refit_top<K>, dim3(n), dim3(kTopThreads), 0, c->stream, ent, c->dtops.p, c->nodes.p, c->dbodies.p, bb, c->dt);
  }
};
// Super levels of at most this many chunks go to the one-block-per-body
// k_refit_top, a wave per chunk: its 16 waves take them in at most two
// passes, against a launch of its own otherwise.  The block's levels together
// hold at most kTopRoots chunks (its LDS table) and kTopMetas internal nodes.
constexpr uint32_t kTopChunks = 32;

// Build the refit plan of a body list into c->plan (device) and c->rl.
// Plan layout (uint32): [bb n][vpre n+1][tpre n+1][partial chunks: body m,
// chunk m] then per super step s: [bb m][gpre m+1][cfirst m][ncnk m]
// [cbase m][cn m], then k_refit_top's entries.
int build_refit_plan(kdop_ctx* c, uint32_t n, const uint32_t* bodies) {
  c->rl.valid = false;
  c->plan_key.clear();
  std::vector<uint32_t> plan;
  plan.reserve(8 * (size_t)n + 16);
  uint32_t maxH = 1;
  for (uint32_t i = 0; i < n; ++i) maxH = std::max(maxH, c->bodies[bodies[i]].H);
  const size_t o_bb = 0;
  for (uint32_t i = 0; i < n; ++i) plan.push_back(bodies[i]);
  // vertices: one thread per vertex, each body's share padded to 64
  const size_t o_vpre = plan.size();
  uint64_t acc = 0;
  plan.push_back(0);
  for (uint32_t i = 0; i < n; ++i) { acc += (c->bodies[bodies[i]].nv + 63ull) / 64 * 64; plan.push_back((uint32_t)acc); }
  const uint64_t total_v = acc;
  // full branch chunks: one thread per tet, each body's range padded to 64;
  // bodies with L = 16 first, their ranges padded to 256 so that each starts
  // on a block (the leaves launch refits their full super chunks of level
  // H-2 per block).  When every body has L = 8 the leaves launch takes
  // 512-thread blocks, each range padded to 512, and a block refits its
  // super chunk of level H-3.  tstart: each item's first thread, then the
  // total.
  bool blk8 = false;
  for (uint32_t i = 0; i < n; ++i) {
    const BodyHost& b = c->bodies[bodies[i]];
    if (b.wide()) continue;
    if (b.L != 8u) { blk8 = false; break; }
    blk8 = true;
  }
  const size_t o_tpre = plan.size();
  plan.insert(plan.end(), (size_t)n + 1, 0u);
  std::vector<uint32_t> order;  // items in the order of their ranges
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t i = 0; i < n; ++i) {
      const BodyHost& b = c->bodies[bodies[i]];
      if (!b.wide() && (b.L == 16u) == (pass == 0)) order.push_back(i);
    }
  acc = 0;
  std::vector<uint32_t> pbody, pchunk;  // partial branch chunks left to k_refit_branch (filled below)
  std::vector<uint32_t> twave;
  for (const uint32_t i : order) {
    const BodyHost& b = c->bodies[bodies[i]];
    const uint64_t full = (uint64_t)(b.nt / b.L) * b.L, gran = blk8 ? 512u : (b.L == 16u ? 256u : 64u);
    plan[o_tpre + i] = (uint32_t)acc;
    const uint64_t len = (full + gran - 1) / gran * gran;
    acc += len;
    twave.insert(twave.end(), (size_t)(len / 64), i);
  }
  plan[o_tpre + n] = (uint32_t)acc;
  const uint64_t total_leaf_threads = acc;
  if (total_leaf_threads >= (1ull << 32)) return fail(c, KDOP_ERR_ARG, "kdop_refit: too many tetrahedra in one batch");
  if (total_v >= (1ull << 32)) return fail(c, KDOP_ERR_ARG, "kdop_refit: too many vertices in one batch");
  // per wave of each range, its batch item
  const size_t o_vwave = plan.size();
  for (uint32_t i = 0; i < n; ++i) plan.insert(plan.end(), (plan[o_vpre + i + 1] - plan[o_vpre + i]) / 64, i);
  const size_t o_twave = plan.size();
  plan.insert(plan.end(), twave.begin(), twave.end());
  // super level h of body b: {slot of its first chunk, chunks, elements of
  // the last chunk, first chunk to refit}; none when the leaves pass did it
  auto super_level = [&](const BodyHost& b, int h, uint4& lv) -> bool {
    // level H-2 chunks over L full branch chunks were refitted with the leaves (L * L <= 64: per wave, L = 16: per
    // block), and with 512-thread leaves blocks (every body L = 8) the full level H-3 chunks too
    const uint32_t cf = (h == (int)b.H - 2 && (b.L * b.L <= 64 || b.L == 16)) ? (b.nt / b.L) / b.L
                        : (h == (int)b.H - 3 && blk8) ? ((b.nt / b.L) / b.L) / b.L
                                                      : 0u;
    const uint32_t nc = b.C[h], nl = host_chunk_elems(b, h, nc - 1);
    if (cf >= nc) return false;
    const uint32_t nfc = (nc - 1 - cf) + (nl == b.L ? 1u : 0u);
    if (!nfc && nl < 2) return false;
    lv = make_uint4(b.node_base + b.level_base[h], nc, nl, cf);
    return true;
  };
  // per body the top super levels (those of at most kTopChunks chunks and
  // every level above them, within the LDS limits) go to k_refit_top; the
  // wider ones below get a launch per step
  std::vector<int> htop(n, -1);
  for (uint32_t i = 0; i < n; ++i) {
    const BodyHost& b = c->bodies[bodies[i]];
    if (b.wide()) continue;
    uint32_t rows = 0, nlev = 0;
    for (int h = 0; h + 2 <= (int)b.H; ++h) {
      uint4 lv;
      const bool has = super_level(b, h, lv);
      const uint32_t r = has ? lv.y - lv.w : 0u;
      if (r > kTopChunks || rows + r > kTopRoots || (rows + r) * (b.L - 1) > kTopMetas || nlev + (has ? 1u : 0u) > kTopEntries - 2) break;
      rows += r;
      nlev += has ? 1u : 0u;
      htop[i] = h;
    }
  }
  // a body's partial branch chunk goes to k_refit_top when that block refits
  // every level above it (level H-2 and up), else to k_refit_branch, before
  // the super launches that need its root
  std::vector<uint32_t> top_branch(n, 0u);  // partial branch chunk + 1 per item
  for (uint32_t i = 0; i < n; ++i) {
    const BodyHost& b = c->bodies[bodies[i]];
    if (b.wide() || b.nt % b.L == 0) continue;
    const uint32_t chunk = b.C[b.H - 1] - 1;
    if (htop[i] >= (int)b.H - 2) top_branch[i] = chunk + 1u;
    else { pbody.push_back(bodies[i]); pchunk.push_back(chunk); }
  }
  const size_t o_pbody = plan.size();
  plan.insert(plan.end(), pbody.begin(), pbody.end());
  const size_t o_pchunk = plan.size();
  plan.insert(plan.end(), pchunk.begin(), pchunk.end());
  using Step = RefitStep;
  std::vector<Step> steps;
  for (uint32_t s = 1; s < maxH; ++s) {
    std::vector<uint32_t> bb, gpre{0}, cfirst, ncnk, cbase, cn;
    uint32_t tot = 0;
    for (uint32_t i = 0; i < n; ++i) {
      const BodyHost& b = c->bodies[bodies[i]];
      if (b.H < s + 1 || b.wide()) continue;
      const int h = (int)b.H - 1 - (int)s;
      if (h <= htop[i]) continue;
      uint4 lv;
      if (!super_level(b, h, lv)) continue;
      const uint32_t cf = lv.w, nc = lv.y, nl = lv.z;
      const uint32_t nfc = (nc - 1 - cf) + (nl == b.L ? 1u : 0u);
      const uint32_t blocks = nfc + (nl < b.L && nl >= 2 ? 1u : 0u);  // a wave per chunk
      bb.push_back(bodies[i]);
      tot += blocks;
      gpre.push_back(tot);
      cfirst.push_back(cf);
      ncnk.push_back(nc);
      cbase.push_back(b.node_base + b.level_base[h]);
      cn.push_back(nl);
    }
    if (bb.empty()) continue;
    Step st;
    st.m = (uint32_t)bb.size();
    st.blocks = tot;
    st.o_bb = plan.size(); plan.insert(plan.end(), bb.begin(), bb.end());
    st.o_gpre = plan.size(); plan.insert(plan.end(), gpre.begin(), gpre.end());
    st.o_gitem = plan.size();  // per block, its item of the step
    for (uint32_t j = 0; j < st.m; ++j) plan.insert(plan.end(), gpre[j + 1] - gpre[j], j);
    st.o_cfirst = plan.size(); plan.insert(plan.end(), cfirst.begin(), cfirst.end());
    st.o_ncnk = plan.size(); plan.insert(plan.end(), ncnk.begin(), ncnk.end());
    st.o_cbase = plan.size(); plan.insert(plan.end(), cbase.begin(), cbase.end());
    st.o_cn = plan.size(); plan.insert(plan.end(), cn.begin(), cn.end());
    steps.push_back(st);
  }
  // k_refit_top's entries: kTopEntries 16-byte records per body, its levels
  // deepest first
  while (plan.size() % 4) plan.push_back(0);
  const size_t o_top = plan.size();
  for (uint32_t i = 0; i < n; ++i) {
    const BodyHost& b = c->bodies[bodies[i]];
    std::vector<uint4> levs;
    uint32_t rows = 0;
    for (int h = htop[i]; h >= 0; --h) {
      uint4 lv;
      if (super_level(b, h, lv)) { levs.push_back(lv); rows += lv.y - lv.w; }
    }
    const uint32_t nmeta = rows * (b.L - 1);
    plan.insert(plan.end(), {(uint32_t)levs.size(), b.L, b.ntop, b.top_base, b.node_base, rows, nmeta, top_branch[i]});
    for (uint32_t e = 0; e < kTopEntries - 2; ++e) {
      const uint4 lv = e < levs.size() ? levs[e] : make_uint4(0, 0, 0, 0);
      plan.insert(plan.end(), {lv.x, lv.y, lv.z, lv.w});
    }
  }
  int rc = ensure(c, c->plan, plan.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->plan.p, plan.data(), plan.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  RefitLaunch& R = c->rl;
  R.o_bb = o_bb; R.o_vpre = o_vpre; R.o_tpre = o_tpre; R.o_vwave = o_vwave; R.o_twave = o_twave; R.o_pbody = o_pbody; R.o_pchunk = o_pchunk; R.o_top = o_top;
  R.total_v = total_v;
  R.total_leaf_threads = total_leaf_threads;
  R.npbody = (uint32_t)pbody.size();
  R.blk8 = blk8;
  R.steps = steps;
  R.wide.clear();
  for (uint32_t i = 0; i < n; ++i)
    if (c->bodies[bodies[i]].wide()) R.wide.push_back(bodies[i]);
  R.valid = true;
  c->plan_key.assign(bodies, bodies + n);
  return KDOP_OK;
}

// Transform (or copy, identity) + leaf/union passes for a batch of bodies.
// pos/quat: device pointers (3n, 4n floats) or null with identity.  The
// plan depends on the body list alone, so a repeated list reuses it.
int run_refit(kdop_ctx* c, uint32_t n, const uint32_t* bodies, const float* pos_dev, const float* quat_dev, bool identity) {
  const bool same = c->rl.valid && c->plan_key.size() == n && std::equal(bodies, bodies + n, c->plan_key.begin());
  if (!same) {
    const int rc = build_refit_plan(c, n, bodies);
    if (rc) return rc;
  }
  const RefitLaunch& R = c->rl;
  const uint32_t* P = c->plan.p;
  // leaves and the vertex transform in one launch, its grid over both ranges
  const uint64_t leaves_grid = std::max<uint64_t>(R.total_leaf_threads, R.total_v);
  if (leaves_grid)
    dispatch_K<RefitLeaves>(c->K, c, (uint32_t)leaves_grid, R.blk8, P + R.o_bb, P + R.o_tpre, P + R.o_vpre, P + R.o_twave, P + R.o_vwave,
                            n, pos_dev, quat_dev, identity);
  if (R.npbody) dispatch_K<RefitBranch>(c->K, c, R.npbody, P + R.o_pbody, P + R.o_pchunk);
  for (const RefitStep& st : R.steps)
    dispatch_K<RefitSuper>(c->K, c, st.blocks, P + st.o_bb, P + st.o_gpre, P + st.o_gitem, P + st.o_cfirst, P + st.o_ncnk, P + st.o_cbase, P + st.o_cn,
                           st.m);
  for (const uint32_t id : R.wide) dispatch_K<RefitWide>(c->K, c, id, c->bodies[id]);
  dispatch_K<RefitTop>(c->K, c, n, reinterpret_cast<const uint4*>(P + R.o_top), P + R.o_bb);
  HIPCHK(c, hipGetLastError());
  if (!c->async_refit || identity) HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

// --------------------------------------------------------------- traversal
template <int K>
struct Seed {
  // expand: write level 1 (the seeds' expansion, into front[1]) instead of level 0
  static void run(kdop_ctx* c, const TravArgs& T, uint32_t total, uint32_t* fill, uint32_t cap, bool expand) {
    const uint32_t blocks = std::max(1u, std::min(4096u, (total + 255) / 256));
    if (expand && c->item_lanes == 4 && K >= 12)  // four lanes per seed pair, like the ramp levels
      hipLaunchKernelGGL(k_seed_quad<K>, dim3(std::max(1u, std::min(4096u, (total + 63) / 64))), dim3(256), 0, c->stream, T, c->dseeds.p,
                         total, c->front[1].p, fill, cap);
    else if (expand)
      hipLaunchKernelGGL((k_seed<K, true>), dim3(blocks), dim3(256), 0, c->stream, T, c->dseeds.p, total, c->front[1].p, fill, cap);
    else
      hipLaunchKernelGGL((k_seed<K, false>), dim3(blocks), dim3(256), 0, c->stream, T, c->dseeds.p, total, c->front[0].p, fill, cap);
  }
};
// A breadth-first level: the persistent strategy's short ramp keeps each level
// one contiguous run (round 0 reads it so); the level-synchronous traversal
// writes every level after the seeds in kSegs segments.
template <int K>
struct Bfs {
  static void run(kdop_ctx* c, const TravArgs& T, uint32_t cap, uint32_t level, uint32_t target, uint32_t last, bool segmented) {
    const uint32_t nin = (segmented && level > 0) ? (uint32_t)kSegs : 1u, nout = segmented ? (uint32_t)kSegs : 1u;
    // four lanes per item (like the persistent rounds) for tight DOPs; loose
    // DOPs (K <= 8), whose levels are wide, run faster with one lane per item
    // (C2 level-synchronous: K=8 0.241 vs 0.251 ms, K=26 0.311 vs 0.260)
    if (c->item_lanes == 4 && K >= 12)
      hipLaunchKernelGGL((k_bfs_level<K, true>), dim3((unsigned)(c->num_cus * 4)), dim3(256), 0, c->stream, T, c->front[0].p, c->front[1].p,
                         cap, c->dstate, level, target, last, nin, nout);
    else
      hipLaunchKernelGGL((k_bfs_level<K, false>), dim3((unsigned)(c->num_cus * 4)), dim3(256), 0, c->stream, T, c->front[0].p,
                         c->front[1].p, cap, c->dstate, level, target, last, nin, nout);
  }
};
// A persistent round, launched with start / stop events attached to its
// dispatch packet (the events time exactly this kernel, with no extra packets
// on the stream).
// Persistent rounds queued before the first host sync: as many as the
// context's recent calls needed (one at first; up to every round a call may
// run, kMaxRounds - 1; raised at once, lowered after 8 calls in a row needed
// fewer), so that a query whose round 0 spills nothing launches no empty
// rounds, and a batch whose round budget spreads heavy pairs over several
// rounds makes no host round trip between them (512-pair batch: 7 rounds,
// each drained round cost a summary and a host sync, 25-40 us).  A call that
// needs more than were queued drains them one by one after a host sync.
// Rounds after the queued ones hand no stack tail on (only a full stack
// spills), so a small stack keep cannot push work from round to round until
// the rounds run out.
constexpr uint32_t kQueuedRounds = kMaxRounds - 1;
// Breadth-first levels before the persistent rounds when the caller set none:
// a call whose BVTT stays narrow gains little balance from a wider frontier
// and pays a launch per level.  Per query (refit, traversal, narrow phase),
// 1 / 2 / 4 / 6 levels: 5k-tet pair K=8 0.099 / 0.104 / 0.117 / 0.104 ms,
// 1k-tet pair 0.080 / 0.087 / 0.106 / 0.094, 40k-tet pair K=8 0.189 / 0.190 /
// 0.162 / 0.164, K=26 0.160 / 0.159 / 0.163 / 0.177; two 320k-tet bodies at
// K=26: 3 / 4 / 5 / 6 / 7 levels 0.189 / 0.189 / 0.189-0.191 / 0.193-0.195 /
// 0.197 ms per step, and 64 pairs of 20k-80k tets at K=18 (a refit series
// step) 4 / 5 / 6 levels 0.883 / 0.868 / 0.878 ms: 5.  call_tets: tets of
// both bodies over the call's pairs.
inline uint32_t bfs_auto(uint64_t call_tets) { return call_tets < 32000 ? 1u : (call_tets < 400000 ? 4u : 5u); }
inline uint32_t round_keep(const kdop_ctx* c, uint32_t round) { return round < c->rounds_queued ? c->keep : kNoHandOffKeep; }

// A persistent round.  With kernel timing on, start / stop events ride on its
// dispatch packet and time exactly this kernel.
template <int K, int LS>
void launch_round(kdop_ctx* c, const TravArgs& T, uint32_t cap, uint32_t round, uint32_t last) {
  const dim3 grid((unsigned)(c->num_cus * wpc_of(c))), block(64);
  Item* b0 = c->front[0].p;
  Item* b1 = c->front[1].p;
  if (c->timing_level >= 2)
    hipExtLaunchKernelGGL((k_persist<K, LS>), grid, block, 0, c->stream, c->evr[2 * round], c->evr[2 * round + 1], 0, T, b0, b1,
                          cap, c->dstate, round, round_keep(c, round), last);
  else
    hipLaunchKernelGGL((k_persist<K, LS>), grid, block, 0, c->stream, T, b0, b1, cap, c->dstate, round, round_keep(c, round), last);
}
// A persistent round of W-wave workgroups sharing one stack of 512 entries per
// wave (k_persist_blk); as many waves as the per-wave kernel launches.
template <int K, int W>
void launch_round_blk(kdop_ctx* c, const TravArgs& T, uint32_t cap, uint32_t round, uint32_t last) {
  const dim3 grid((unsigned)std::max(1, c->num_cus * wpc_of(c) / W)), block(64 * W);
  Item* b0 = c->front[0].p;
  Item* b1 = c->front[1].p;
  if (c->timing_level >= 2)
    hipExtLaunchKernelGGL((k_persist_blk<K, 512, W>), grid, block, 0, c->stream, c->evr[2 * round], c->evr[2 * round + 1], 0, T, b0,
                          b1, cap, c->dstate, round, round_keep(c, round), last);
  else
    hipLaunchKernelGGL((k_persist_blk<K, 512, W>), grid, block, 0, c->stream, T, b0, b1, cap, c->dstate, round, round_keep(c, round),
                       last);
}
// A persistent round of W-wave workgroups with 4 lanes per work item
// (k_persist_quad) and a shared stack of 2^SL items.
template <int K, int W, int SL, bool COUNT = false, int IPQ = 1, bool SLAB = false>
void launch_round_quad(kdop_ctx* c, const TravArgs& T, uint32_t cap, uint32_t round, uint32_t last) {
  int per_cu = wpc_of(c) / W;
  if (SLAB) {  // the node slabs' LDS holds fewer workgroups per CU: launch as many as are resident
    static int occ = -1;
    if (occ < 0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_persist_quad<K, W, SL, COUNT, IPQ, SLAB>, 64 * W, 0) != hipSuccess)
      occ = 1;
    per_cu = std::min(per_cu, std::max(1, occ));
  }
  const dim3 grid((unsigned)std::max(1, c->num_cus * per_cu)), block(64 * W);
  Item* b0 = c->front[0].p;
  Item* b1 = c->front[1].p;
  // the budget applies in rounds that may spill (the last one allowed takes
  // no budget: its spill would only go to the level-synchronous finish)
  const uint32_t budget = (c->round_iters_eff && round + 1 < c->max_rounds) ? c->round_iters_eff : ~0u;
  if (c->timing_level >= 2)
    hipExtLaunchKernelGGL((k_persist_quad<K, W, SL, COUNT, IPQ, SLAB>), grid, block, 0, c->stream, c->evr[2 * round], c->evr[2 * round + 1],
                          0, T, b0, b1, cap, c->dstate, round, round_keep(c, round), last, budget);
  else
    hipLaunchKernelGGL((k_persist_quad<K, W, SL, COUNT, IPQ, SLAB>), grid, block, 0, c->stream, T, b0, b1, cap, c->dstate, round,
                       round_keep(c, round), last, budget);
}
template <int K>
struct PersistRound {
  static void run(kdop_ctx* c, const TravArgs& T, uint32_t cap, uint32_t round, uint32_t last) {
    const int bw = bw_of(c), ls = ls_of(c);
    if (c->count_records) {  // record-counting calls: the quad kernel in its default shape
      launch_round_quad<K, 4, 9, true>(c, T, cap, round, last);
      return;
    }
    if (c->item_lanes == 4 && c->node_slab) {  // records staged in per-wave LDS node slabs
      if (bw <= 2) launch_round_quad<K, 2, 9, false, 1, true>(c, T, cap, round, last);
      else launch_round_quad<K, 4, 10, false, 1, true>(c, T, cap, round, last);
      return;
    }
    if (c->item_lanes == 4 && c->quad_items == 2) {  // two items per quad, 4-wave workgroups
      if (ls >= 512) launch_round_quad<K, 4, 11, false, 2>(c, T, cap, round, last);
      else launch_round_quad<K, 4, 10, false, 2>(c, T, cap, round, last);
      return;
    }
    if (c->item_lanes == 4) {  // stack of KDOP_OPT_QUEUE_DEPTH (256 or 512) entries per wave
      const bool deep = ls >= 512;
      if (bw <= 2) { if (deep) launch_round_quad<K, 2, 10>(c, T, cap, round, last); else launch_round_quad<K, 2, 9>(c, T, cap, round, last); }
      else if (bw == 4) {
        if (deep) launch_round_quad<K, 4, 11>(c, T, cap, round, last);
        else if (ls <= 128) launch_round_quad<K, 4, 9>(c, T, cap, round, last);
        else launch_round_quad<K, 4, 10>(c, T, cap, round, last);
      }
      else launch_round_quad<K, 8, 11>(c, T, cap, round, last);
      return;
    }
    if (bw == 2) { launch_round_blk<K, 2>(c, T, cap, round, last); return; }
    if (bw == 4) { launch_round_blk<K, 4>(c, T, cap, round, last); return; }
    if (bw == 8) { launch_round_blk<K, 8>(c, T, cap, round, last); return; }
    switch (ls) {
      case 256: launch_round<K, 256>(c, T, cap, round, last); break;
      case 512: launch_round<K, 512>(c, T, cap, round, last); break;
      case 2048: launch_round<K, 2048>(c, T, cap, round, last); break;
      default: launch_round<K, 1024>(c, T, cap, round, last); break;
    }
  }
};

// Device scratch of a PROX reduction of n records: two key arrays (the keys
// in the first), keep flags, their offsets, the scan's block sums and the
// kept count.
struct ReduceScratch {
  ProxKey *ka, *kb;
  uint32_t *keep, *off, *sums, *total;
};
// ------------------------------------------------ allocation-free reduction
// The scratch of kdop_reduce_prox_contacts_noalloc: one fixed array per GPU,
// held until the process ends, laid out for KDOP_NOALLOC_MAX_CONTACTS
// records (two key arrays, keep flags, offsets, block sums, the count).
constexpr size_t kNaKeyBytes = (size_t)KDOP_NOALLOC_MAX_CONTACTS * sizeof(ProxKey);
constexpr size_t kNaWords = 2 * (size_t)KDOP_NOALLOC_MAX_CONTACTS + (KDOP_NOALLOC_MAX_CONTACTS + kScanBlock - 1) / kScanBlock + 1;
constexpr size_t kNoAllocScratchBytes = 2 * kNaKeyBytes + kNaWords * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on each GPU (the host-side name is not one),
// looked up once per device with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

// Called by kdop_create with the context's device current.  A failed lookup
// leaves ctx->scratch null and its message in ctx->scratch_err; only the
// allocation-free entry point depends on it.
void setup_noalloc_scratch(kdop_ctx* c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  const size_t dev = (size_t)c->cfg.device;
  if (g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    g_scratch_ptr[dev] = p;
  }
  c->scratch = g_scratch_ptr[dev];
}

ReduceScratch noalloc_scratch_of(void* base) {
  ReduceScratch r;
  unsigned char* b = static_cast<unsigned char*>(base);
  r.ka = reinterpret_cast<ProxKey*>(b);
  r.kb = reinterpret_cast<ProxKey*>(b + kNaKeyBytes);
  r.keep = reinterpret_cast<uint32_t*>(b + 2 * kNaKeyBytes);
  r.off = r.keep + KDOP_NOALLOC_MAX_CONTACTS;
  r.sums = r.off + KDOP_NOALLOC_MAX_CONTACTS;
  r.total = r.sums + (KDOP_NOALLOC_MAX_CONTACTS + kScanBlock - 1) / kScanBlock;
  return r;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void kdop_default_config(kdop_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->K = 8;
  cfg->device = 0;
  cfg->chunk_bytes = 8000;
  cfg->traversal = KDOP_TRAVERSAL_AUTO;
}

int kdop_create(kdop_ctx** out, const kdop_config* cfg) {
  if (!out) return KDOP_ERR_ARG;
  *out = nullptr;
  kdop_config def;
  kdop_default_config(&def);
  if (!cfg) cfg = &def;
  if (!valid_K(cfg->K)) return KDOP_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KDOP_ERR_HIP;
  if (cfg->device < 0 || cfg->device >= ndev) return KDOP_ERR_ARG;
  if (hipSetDevice(cfg->device) != hipSuccess) return KDOP_ERR_HIP;
  kdop_ctx* c = new kdop_ctx();
  c->cfg = *cfg;
  c->K = cfg->K;
  make_dir_table(c->K, c->dt);
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return KDOP_ERR_HIP; }
  c->own_stream = true;
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_HIP; }
  for (auto& e : c->evr)
    if (hipEventCreate(&e) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_HIP; }
  if (hipMalloc(&c->dstate, sizeof(CallState)) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_NOMEM; }
  if (hipHostMalloc(&c->hsum, sizeof(CallSummary), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    kdop_destroy(c);
    return KDOP_ERR_NOMEM;
  }
  std::memset(c->hsum, 0, sizeof(CallSummary));
  if (hipHostGetDevicePointer((void**)&c->hsum_dev, c->hsum, 0) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_HIP; }
  if (hipHostMalloc(&c->hwords, sizeof(HostWords), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    kdop_destroy(c);
    return KDOP_ERR_NOMEM;
  }
  std::memset(c->hwords, 0, sizeof(HostWords));
  if (hipHostGetDevicePointer((void**)&c->hwords_dev, c->hwords, 0) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_HIP; }
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess) { kdop_destroy(c); return KDOP_ERR_HIP; }
    c->num_cus = prop.multiProcessorCount;
  }
  setup_noalloc_scratch(c);
  {
    const int64_t opts[5][2] = {{KDOP_OPT_QUEUE_DEPTH, cfg->queue_depth}, {KDOP_OPT_WAVES_PER_CU, cfg->waves_per_cu},
                                {KDOP_OPT_RAMP_TARGET, cfg->ramp_target}, {KDOP_OPT_STACK_KEEP, cfg->stack_keep},
                                {KDOP_OPT_SEED_BUDGET, cfg->seed_budget}};
    for (const auto& o : opts)
      if (o[1] != 0 && kdop_set_option(c, (int)o[0], o[1]) != KDOP_OK) { kdop_destroy(c); return KDOP_ERR_ARG; }
  }
  *out = c;
  return KDOP_OK;
}

void kdop_destroy(kdop_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->bodies) {
    (void)hipFree(b.X0); (void)hipFree(b.Y0); (void)hipFree(b.Z0); (void)hipFree(b.W0);
    (void)hipFree(b.X); (void)hipFree(b.Y); (void)hipFree(b.Z); (void)hipFree(b.W);
    (void)hipFree(b.tets); (void)hipFree(b.surf);
    if (b.leaf_slot) (void)hipFree(b.leaf_slot);
    if (b.unions) (void)hipFree(b.unions);
  }
  release(c->nodes); release(c->dbodies); release(c->dposes); release(c->dtops); release(c->dpairs); release(c->dseeds);
  release(c->front[0]); release(c->front[1]); release(c->p1); release(c->p2);
  release(c->contacts); release(c->recs); release(c->gate); release(c->adj_acc); release(c->adj_bodies); release(c->plan); release(c->poses);
  release(c->contacts_tmp); release(c->sortbuf); release(c->raybuf); release(c->sphere_out);
  release(c->redbuf); release(c->prox_io); release(c->sg_items[0]); release(c->sg_items[1]); release(c->sg_scan); release(c->sg_kind);
  release(c->rc_best); release(c->rc_first); release(c->touched);
  if (c->dstate) (void)hipFree(c->dstate);
  if (c->hsum) (void)hipHostFree(c->hsum);
  if (c->hwords) (void)hipHostFree(c->hwords);
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  for (auto& e : c->evr) if (e) (void)hipEventDestroy(e);
  if (c->tg_exec) (void)hipGraphExecDestroy(c->tg_exec);
  if (c->tg) (void)hipGraphDestroy(c->tg);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* kdop_last_error(const kdop_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* kdop_stream(kdop_ctx* c) { return c ? (void*)c->stream : nullptr; }

int kdop_set_option(kdop_ctx* c, int option, int64_t v) {
  if (!c) return KDOP_ERR_ARG;
  switch (option) {
    case KDOP_OPT_TRAVERSAL:
      if (v < 0 || v > 2) break;
      c->cfg.traversal = (int32_t)v;
      return KDOP_OK;
    case KDOP_OPT_QUEUE_DEPTH:
      if (v != 0 && v != 128 && v != 256 && v != 512 && v != 1024 && v != 2048) break;
      c->ls = (int)v;
      return KDOP_OK;
    case KDOP_OPT_WAVES_PER_CU:
      if (v < 0 || v > 32) break;
      c->wpc = (int)v;
      return KDOP_OK;
    case KDOP_OPT_RAMP_TARGET:
      if (v < 0 || v > (1ll << 30)) break;
      c->ramp_target = (uint32_t)v;
      return KDOP_OK;
    case KDOP_OPT_STACK_KEEP:
      if (v < 0 || v > (1ll << 20)) break;
      c->keep = v == 0 ? kNoHandOffKeep : (uint32_t)v;
      return KDOP_OK;
    case KDOP_OPT_SEED_BUDGET:  // 0: by the call's pair count
      if (v < 0 || v > (1ll << 30)) break;
      c->seed_budget = v ? (uint32_t)v : kSeedBudgetDefault;
      c->seed_budget_set = v != 0;
      return KDOP_OK;
    case KDOP_OPT_KERNEL_TIMING:
      if (v < 0 || v > 2) break;
      c->timing_level = (int)v;
      return KDOP_OK;
    case KDOP_OPT_GRAPHS:
      if (v != 0 && v != 1) break;
      c->use_graphs = v != 0;
      return KDOP_OK;
    case KDOP_OPT_NARROW_PATH:
      if (v < 0 || v > 2) break;
      c->narrow_path = (int)v;
      return KDOP_OK;
    case KDOP_OPT_ASYNC_REFIT:
      if (v != 0 && v != 1) break;
      c->async_refit = v != 0;
      return KDOP_OK;
    case KDOP_OPT_BFS_MAX:
      if (v < 0 || v > 62) break;
      c->bfs_max = (uint32_t)v;  // 0: by the call's size
      return KDOP_OK;
    case KDOP_OPT_BLOCK_WAVES:
      if (v != 0 && v != 1 && v != 2 && v != 4 && v != 8) break;
      c->bw = (int)v;
      return KDOP_OK;
    case KDOP_OPT_ROUND_ITERS:  // 0: no budget
      if (v < 0 || v > (1ll << 20)) break;
      c->round_iters = (uint32_t)v;
      c->round_iters_set = true;
      return KDOP_OK;
    case KDOP_OPT_QUAD_ITEMS:
      if (v == 0) v = 1;
      if (v != 1 && v != 2) break;
      c->quad_items = (int)v;
      return KDOP_OK;
    case KDOP_OPT_NODE_SLAB:
      if (v != 0 && v != 1) break;
      c->node_slab = (int)v;
      return KDOP_OK;
    case KDOP_OPT_COUNT_RECORDS:
      if (v != 0 && v != 1) break;
      c->count_records = v != 0;
      return KDOP_OK;
    case KDOP_OPT_MAX_ROUNDS:
      if (v == 0) v = kMaxRounds - 1;
      if (v < 1 || v > kMaxRounds - 1) break;
      c->max_rounds = (uint32_t)v;
      return KDOP_OK;
    case KDOP_OPT_ITEM_LANES:
      if (v == 0) v = 4;
      if (v != 1 && v != 4) break;
      c->item_lanes = (int)v;
      return KDOP_OK;
    case KDOP_OPT_SINGLE_ITEMS:
      if (v == 0) v = 1ll << 28;
      if (v < 1 || v > (1ll << 30)) break;
      c->sg_max_items = (uint32_t)v;
      return KDOP_OK;
    default:
      return fail(c, KDOP_ERR_ARG, "kdop_set_option: unknown option %d", option);
  }
  return fail(c, KDOP_ERR_ARG, "kdop_set_option: value %lld out of range for option %d", (long long)v, option);
}

int kdop_autotune(kdop_ctx* c, uint32_t npairs, const uint32_t* body_pairs, uint32_t flags, double budget_ms, int64_t* chosen,
                  double* best_ms) {
  if (!c) return KDOP_ERR_ARG;
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  auto elapsed_ms = [&]() { return std::chrono::duration<double, std::milli>(clk::now() - t_start).count(); };
  // median wall time of 5 calls with the context's current options
  auto measure = [&](double* ms) -> int {
    double t[5];
    for (int r = 0; r < 5; ++r) {
      const auto a = clk::now();
      const int rc = kdop_traverse(c, npairs, body_pairs, flags, nullptr);
      if (rc) return rc;
      t[r] = std::chrono::duration<double, std::milli>(clk::now() - a).count();
    }
    std::sort(t, t + 5);
    *ms = t[2];
    return KDOP_OK;
  };
  int rc = kdop_traverse(c, npairs, body_pairs, flags, nullptr);  // warm up (buffers sized)
  if (rc) return rc;
  double best = 0.0;
  if ((rc = measure(&best))) return rc;
  const bool start_persistent = c->last_persistent;  // the strategy the starting options ran (AUTO resolved)
  // candidate values per option; the launch-shape ones depend on the lanes
  // per item chosen before them (the quad kernel runs more, wider workgroups)
  auto candidates = [&](int opt) -> std::vector<int64_t> {
    const bool quad = c->item_lanes == 4;
    const int64_t pw = (int64_t)c->num_cus * wpc_of(c);  // persistent waves in effect
    const int64_t ls = ls_of(c);
    switch (opt) {
      case KDOP_OPT_TRAVERSAL: return {KDOP_TRAVERSAL_PERSISTENT, KDOP_TRAVERSAL_LEVELS};
      case KDOP_OPT_ITEM_LANES: return {4, 1};
      case KDOP_OPT_BLOCK_WAVES: return quad ? std::vector<int64_t>{2, 4, 8} : std::vector<int64_t>{1, 2, 4};
      case KDOP_OPT_WAVES_PER_CU: return quad ? std::vector<int64_t>{12, 16, 24} : std::vector<int64_t>{4, 8, 12};
      case KDOP_OPT_QUEUE_DEPTH: return quad ? std::vector<int64_t>{128, 256, 512} : std::vector<int64_t>{256, 512, 1024};
      // no hand-off, or a keep that hands on the bottom of deep stacks
      case KDOP_OPT_STACK_KEEP: return {0, ls / 2, ls / 8};
      case KDOP_OPT_ROUND_ITERS: return {0, 16, 24, 32, 64};
      case KDOP_OPT_NODE_SLAB: return quad ? std::vector<int64_t>{0, 1} : std::vector<int64_t>{0};
      case KDOP_OPT_RAMP_TARGET: return {4 * pw, 16 * pw, 64 * pw};
      default: return {16384, 65536, 262144};  // KDOP_OPT_SEED_BUDGET
    }
  };
  const int order[] = {KDOP_OPT_TRAVERSAL, KDOP_OPT_ITEM_LANES, KDOP_OPT_NODE_SLAB, KDOP_OPT_ROUND_ITERS, KDOP_OPT_BLOCK_WAVES, KDOP_OPT_WAVES_PER_CU,
                       KDOP_OPT_QUEUE_DEPTH, KDOP_OPT_STACK_KEEP, KDOP_OPT_RAMP_TARGET, KDOP_OPT_SEED_BUDGET};
  for (const int opt : order) {
    struct {
      int opt;
      std::vector<int64_t> values;
    } k{opt, candidates(opt)};
    int64_t cur = 0;
    kdop_get_option(c, k.opt, &cur);
    int64_t best_v = cur;
    for (const int64_t v : k.values) {
      if (v == cur) continue;
      if (elapsed_ms() > budget_ms) break;
      if (kdop_set_option(c, k.opt, v) != KDOP_OK) continue;
      double ms = 0.0;
      rc = measure(&ms);
      if (rc == KDOP_ERR_CAPACITY) { rc = KDOP_OK; continue; }  // a candidate the call cannot run with: skip it
      if (rc) { kdop_set_option(c, k.opt, cur); return rc; }
      // a candidate must beat the best by more than the calls' noise (1 %)
      if (ms < 0.99 * best) { best = ms; best_v = v; }
    }
    // an AUTO start that stayed best is pinned to the strategy it ran
    if (k.opt == KDOP_OPT_TRAVERSAL && best_v == KDOP_TRAVERSAL_AUTO)
      best_v = start_persistent ? KDOP_TRAVERSAL_PERSISTENT : KDOP_TRAVERSAL_LEVELS;
    kdop_set_option(c, k.opt, best_v);
    // the levels strategy ignores the persistent knobs
    if (k.opt == KDOP_OPT_TRAVERSAL && best_v == KDOP_TRAVERSAL_LEVELS) break;
  }
  if ((rc = kdop_traverse(c, npairs, body_pairs, flags, nullptr))) return rc;  // results of the chosen options
  if (chosen)
    for (int o = 1; o <= KDOP_AUTOTUNE_CHOSEN; ++o)  // every option by number (0 for an unknown number)
      if (kdop_get_option(c, o, &chosen[o - 1]) != KDOP_OK) chosen[o - 1] = 0;
  c->err.clear();
  if (best_ms) *best_ms = best;
  return KDOP_OK;
}

int kdop_get_option(kdop_ctx* c, int option, int64_t* v) {
  if (!c || !v) return KDOP_ERR_ARG;
  switch (option) {
    case KDOP_OPT_TRAVERSAL: *v = c->cfg.traversal; return KDOP_OK;
    case KDOP_OPT_QUEUE_DEPTH: *v = ls_of(c); return KDOP_OK;
    case KDOP_OPT_WAVES_PER_CU: *v = wpc_of(c); return KDOP_OK;
    case KDOP_OPT_RAMP_TARGET: *v = c->ramp_target; return KDOP_OK;
    case KDOP_OPT_STACK_KEEP: *v = c->keep >= kNoHandOffKeep ? 0 : c->keep; return KDOP_OK;
    case KDOP_OPT_SEED_BUDGET: *v = c->seed_budget_set ? c->seed_budget : c->seed_budget_eff; return KDOP_OK;
    case KDOP_OPT_KERNEL_TIMING: *v = c->timing_level; return KDOP_OK;
    case KDOP_OPT_ASYNC_REFIT: *v = c->async_refit ? 1 : 0; return KDOP_OK;
    case KDOP_OPT_NARROW_PATH: *v = c->narrow_path; return KDOP_OK;
    case KDOP_OPT_GRAPHS: *v = c->use_graphs ? 1 : 0; return KDOP_OK;
    case KDOP_OPT_BFS_MAX: *v = c->bfs_max; return KDOP_OK;
    case KDOP_OPT_SINGLE_ITEMS: *v = c->sg_max_items; return KDOP_OK;
    case KDOP_OPT_BLOCK_WAVES: *v = bw_of(c); return KDOP_OK;
    case KDOP_OPT_ITEM_LANES: *v = c->item_lanes; return KDOP_OK;
    case KDOP_OPT_MAX_ROUNDS: *v = c->max_rounds; return KDOP_OK;
    case KDOP_OPT_COUNT_RECORDS: *v = c->count_records ? 1 : 0; return KDOP_OK;
    case KDOP_OPT_QUAD_ITEMS: *v = c->quad_items; return KDOP_OK;
    case KDOP_OPT_NODE_SLAB: *v = c->node_slab; return KDOP_OK;
    case KDOP_OPT_ROUND_ITERS: *v = c->round_iters_set ? c->round_iters : c->round_iters_eff; return KDOP_OK;
    default: return fail(c, KDOP_ERR_ARG, "kdop_get_option: unknown option %d", option);
  }
}

int kdop_tree_shape(int32_t K, uint64_t chunk_bytes, uint64_t nt, uint64_t* out) {
  Shape s;
  if (!valid_K(K) || !chunk_shape(K, chunk_bytes, nt, s)) return -1;
  out[0] = s.node_bytes; out[1] = s.n_perfect; out[2] = s.L; out[3] = s.C; out[4] = s.H;
  return 0;
}

int kdop_add_body(kdop_ctx* c, uint32_t nv, const float* X0, const float* Y0, const float* Z0, uint32_t nt, const uint32_t* tets,
                  const uint8_t* surface, uint32_t* body_id) {
  if (!c) return KDOP_ERR_ARG;
  return kdop_add_body_ex(c, c->cfg.chunk_bytes, nv, X0, Y0, Z0, nt, tets, surface, body_id);
}

int kdop_add_body_ex(kdop_ctx* c, uint64_t mem_bytes, uint32_t nv, const float* X0, const float* Y0, const float* Z0, uint32_t nt,
                     const uint32_t* tets, const uint8_t* surface, uint32_t* body_id) {
  if (!c) return KDOP_ERR_ARG;
  if (!X0 || !Y0 || !Z0 || !tets || !surface || !body_id) return fail(c, KDOP_ERR_ARG, "kdop_add_body: null argument");
  if (nv == 0 || nt == 0) return fail(c, KDOP_ERR_ARG, "kdop_add_body: empty mesh (nv=%u nt=%u)", nv, nt);
  if (nt >= (1u << 30)) return fail(c, KDOP_ERR_ARG, "kdop_add_body: too many tetrahedra (%u)", nt);
  for (size_t i = 0; i < 4 * (size_t)nt; ++i)
    if (tets[i] >= nv) return fail(c, KDOP_ERR_ARG, "kdop_add_body: tet vertex index %u out of range (nv=%u)", tets[i], nv);
  Shape s;
  if (!chunk_shape(c->K, mem_bytes, nt, s))
    return fail(c, KDOP_ERR_ARG, "kdop_add_body: mem_bytes=%llu holds no two-leaf chunk for K=%d", (unsigned long long)mem_bytes, c->K);
  if (s.H > (uint64_t)kMaxLevels) return fail(c, KDOP_ERR_ARG, "kdop_add_body: %llu chunk levels", (unsigned long long)s.H);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  BodyHost b;
  b.nv = nv; b.nt = nt; b.L = (uint32_t)s.L; b.H = (uint32_t)s.H;
  b.C.assign(b.H, 0); b.E.assign(b.H, 0); b.level_base.assign(b.H, 0);
  b.C[b.H - 1] = (uint32_t)s.C;
  b.E[b.H - 1] = nt;
  for (int h = (int)b.H - 1; h >= 1; --h) {
    b.C[h - 1] = (uint32_t)chunks_above(b.C[h], b.L);
    b.E[h - 1] = b.C[h];
  }
  // slots: root record (+ a free slot), then level by level; a chunk is one
  // free slot + its nodes, every level starts on an even slot
  uint64_t off = 2;
  for (uint32_t h = 0; h < b.H; ++h) {
    b.level_base[h] = (uint32_t)off;
    const uint32_t nlast = host_chunk_elems(b, (int)h, b.C[h] - 1);
    const bool branch = (h == b.H - 1);
    const uint64_t last_slots = branch ? 2ull * nlast : (nlast >= 2 ? nlast : 0);
    off += (uint64_t)(b.C[h] - 1) * host_stride(b, (int)h) + last_slots;
    b.real_nodes += (uint64_t)(b.C[h] - 1) * (host_stride(b, (int)h) - 1) + (last_slots ? last_slots - 1 : 0);
    off += off & 1u;
  }
  if (c->nodes_used + off >= (1ull << 30)) return fail(c, KDOP_ERR_ARG, "kdop_add_body: node pool exceeds 2^30 nodes");
  b.nnodes = (uint32_t)off;
  c->nodes_used += c->nodes_used & 1u;
  b.node_base = (uint32_t)c->nodes_used;
  uint32_t lg = 0;
  while ((1u << lg) < b.L) ++lg;
  b.depth = 1 + b.H * std::max(lg, 1u);
  // top-level resolved roots
  // resolved chunk roots of every level (level 0 first: the refit's root union)
  b.top_base = (uint32_t)c->tops.size();
  b.roots_off.assign(b.H, 0);
  for (uint32_t h = 0; h < b.H; ++h) {
    b.roots_off[h] = (uint32_t)c->tops.size();
    for (uint32_t t = 0; t < b.C[h]; ++t) c->tops.push_back(host_resolved_root(b, (int)h, t));
  }
  b.ntop = b.C[0];
  // device buffers
  const size_t S = (size_t)rec_stride(c->K);
  int rc = ensure(c, c->nodes, (c->nodes_used + b.nnodes) * S, true);
  if (rc) { c->tops.resize(b.top_base); return rc; }
  auto alloc = [&](void** p, size_t bytes) -> int {
    hipError_t e = hipMalloc(p, bytes);
    return e == hipSuccess ? KDOP_OK : fail(c, KDOP_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
  };
  if ((rc = alloc((void**)&b.X0, nv * 4)) || (rc = alloc((void**)&b.Y0, nv * 4)) || (rc = alloc((void**)&b.Z0, nv * 4)) ||
      (rc = alloc((void**)&b.X, nv * 4)) || (rc = alloc((void**)&b.Y, nv * 4)) || (rc = alloc((void**)&b.Z, nv * 4)) ||
      (rc = alloc((void**)&b.W, nv * (size_t)16)) || (rc = alloc((void**)&b.W0, nv * (size_t)16)) ||
      (rc = alloc((void**)&b.tets, 16 * (size_t)nt)) || (rc = alloc((void**)&b.surf, nt))) {
    c->tops.resize(b.top_base);
    return rc;
  }
  HIPCHK(c, hipMemcpyAsync(b.X0, X0, nv * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.Y0, Y0, nv * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.Z0, Z0, nv * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_interleave3, dim3((nv + 255) / 256), dim3(256), 0, c->stream, (const float*)b.X0, (const float*)b.Y0,
                     (const float*)b.Z0, nv, b.W0);
  HIPCHK(c, hipMemcpyAsync(b.tets, tets, 16 * (size_t)nt, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.surf, surface, nt, hipMemcpyHostToDevice, c->stream));
  const uint32_t id = (uint32_t)c->bodies.size();
  std::vector<uint32_t> hmeta, hleaf, hunions;
  if (b.wide()) {  // the topology from the host, and the refit lists
    build_wide_topology(b, id, surface, hmeta, hleaf, hunions, b.union_off);
    if ((rc = alloc((void**)&b.leaf_slot, 4 * (size_t)nt)) || (rc = alloc((void**)&b.unions, 4 * std::max<size_t>(hunions.size(), 1)))) {
      c->tops.resize(b.top_base);
      return rc;
    }
    HIPCHK(c, hipMemcpyAsync(b.leaf_slot, hleaf.data(), 4 * (size_t)nt, hipMemcpyHostToDevice, c->stream));
    if (!hunions.empty()) HIPCHK(c, hipMemcpyAsync(b.unions, hunions.data(), 4 * hunions.size(), hipMemcpyHostToDevice, c->stream));
  }
  c->nodes_used += b.nnodes;
  c->bodies.push_back(b);
  rc = upload_bodies(c);
  if (rc) return rc;
  // free slots read as meta 0xFFFFFFFF; the root record's meta is 0
  HIPCHK(c, hipMemsetAsync(c->nodes.p + (size_t)b.node_base * S, 0xFF, (size_t)b.nnodes * S * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(c->nodes.p + (size_t)b.node_base * S, 0, S * sizeof(float), c->stream));
  if (b.wide()) {
    uint32_t* dmeta = nullptr;
    if ((rc = alloc((void**)&dmeta, hmeta.size() * 4))) return rc;
    HIPCHK(c, hipMemcpyAsync(dmeta, hmeta.data(), hmeta.size() * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_write_meta, dim3((b.nnodes + 255) / 256), dim3(256), 0, c->stream, (const uint32_t*)dmeta, b.nnodes, b.node_base,
                       c->K, c->nodes.p);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(dmeta);
    if (e != hipSuccess) return fail(c, KDOP_ERR_HIP, "kdop_add_body: topology upload: %s", hipGetErrorString(e));
  } else {  // device construction of the topology
    TopoParams P;
    std::memset(&P, 0, sizeof(P));
    P.node_base = b.node_base; P.H = b.H; P.L = b.L; P.nt = nt; P.body = id;
    uint32_t total_chunks = 0;
    for (uint32_t h = 0; h < b.H; ++h) { P.C[h] = b.C[h]; P.E[h] = b.E[h]; P.level_base[h] = b.level_base[h]; total_chunks += b.C[h]; }
    hipLaunchKernelGGL(k_build_topology, dim3((total_chunks + 63) / 64), dim3(64), 0, c->stream, P, c->nodes.p, c->K, b.surf, total_chunks);
  }
  HIPCHK(c, hipGetLastError());
  rc = run_refit(c, 1, &id, nullptr, nullptr, true);
  if (rc) return rc;
  *body_id = id;
  return KDOP_OK;
}

int kdop_body_shape(kdop_ctx* c, uint32_t body, uint64_t* out) {
  if (!c) return KDOP_ERR_ARG;
  if (body >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "unknown body %u", body);
  const BodyHost& b = c->bodies[body];
  out[0] = b.L; out[1] = b.C[b.H - 1]; out[2] = b.H; out[3] = b.real_nodes; out[4] = b.depth;
  return KDOP_OK;
}

int kdop_refit_device(kdop_ctx* c, uint32_t n, const uint32_t* bodies, const float* pos_dev, const float* quat_dev) {
  if (!c) return KDOP_ERR_ARG;
  if (n == 0) return KDOP_OK;
  if (!bodies || !pos_dev || !quat_dev) return fail(c, KDOP_ERR_ARG, "kdop_refit: null argument");
  for (uint32_t i = 0; i < n; ++i)
    if (bodies[i] >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "kdop_refit: unknown body %u", bodies[i]);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  return run_refit(c, n, bodies, pos_dev, quat_dev, false);
}

int kdop_refit(kdop_ctx* c, uint32_t n, const uint32_t* bodies, const float* pos, const float* quat) {
  if (!c) return KDOP_ERR_ARG;
  if (n == 0) return KDOP_OK;
  if (!bodies || !pos || !quat) return fail(c, KDOP_ERR_ARG, "kdop_refit: null argument");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  int rc = ensure(c, c->poses, 7 * (size_t)n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->poses.p, pos, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->poses.p + 3 * (size_t)n, quat, 4 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  return kdop_refit_device(c, n, bodies, c->poses.p, c->poses.p + 3 * (size_t)n);
}

}  // extern "C"

namespace {

// Narrow-phase arguments; with pending_round >= 0 the gate runs only when
// the persistent traversal left nothing in that round's segments, with
// levels_done only when the level-synchronous traversal finished.
NarrowArgs narrow_args(kdop_ctx* c, const TravArgs& T, int pending_round, bool levels_done) {
  CallState* d = c->dstate;
  NarrowArgs A;
  A.bodies = c->dbodies.p; A.pairs = c->dpairs.p; A.p2 = c->p2.p; A.p2_fill = &d->fills.p2; A.p2_cap = T.p2_cap;
  A.contacts = c->contacts.p; A.c_fill = &d->fills.c; A.c_cap = (uint32_t)std::min<size_t>(c->contacts.cap, 0xFFFFFFFFu);
  A.recs = c->recs.p; A.r_fill = &d->fills.r; A.r_cap = (uint32_t)std::min<size_t>(c->recs.cap, 0xFFFFFFFFu);
  A.stripes = d->stripe;
  A.overflow = &d->overflow;
  A.pending = pending_round >= 0 ? &d->seg[pending_round][0].fill : nullptr;
  A.pending_stride = sizeof(SegCtr) / sizeof(uint32_t);
  A.done = levels_done ? &d->final_set : nullptr;
  A.gate = c->gate.p;
  A.g_fill = &d->fills.g;
  return A;
}

// The narrow phase.  Few surface pairs (as many as the GPU holds waves): one
// wave per pair in one pass; many: the SAT gate with 8 lanes per pair, then
// one wave per surviving pair (most do not survive).  The choice follows the
// previous call's surface-pair count (or the size estimate on the first
// call); both paths give the same results.
bool narrow_one_pass(const kdop_ctx* c, uint64_t p2_guess) {
  return c->narrow_path == 1 || (c->narrow_path == 0 && p2_guess <= (uint64_t)c->num_cus * 32);
}
int launch_narrow(kdop_ctx* c, const TravArgs& T, int pending_round, bool levels_done, bool one_pass) {
  const NarrowArgs A = narrow_args(c, T, pending_round, levels_done);
  // grid-stride kernels whose iterations are chains of dependent gathers: as
  // many resident blocks as their LDS allows (about 13 KB for the gate, 26 KB
  // with the contact stage) hide the chains' latency
  const dim3 block(256);
  if (one_pass) {
    hipLaunchKernelGGL(k_narrow_wave, dim3((unsigned)(c->num_cus * 4)), block, 0, c->stream, A);
  } else {
    hipLaunchKernelGGL(k_narrow_gate, dim3((unsigned)(c->num_cus * 12)), block, 0, c->stream, A);
    hipLaunchKernelGGL(k_narrow_gated, dim3((unsigned)(c->num_cus * 6)), block, 0, c->stream, A);
  }
  HIPCHK(c, hipGetLastError());
  return KDOP_OK;
}

// Fold the call's counters on the device and copy the summary to the host.
// final_round >= 0: the fold also zeroes the call state (for the next call,
// in place of a memset) when the call is complete: no overflow and nothing
// left in that round's segments (-1: levels strategy, no rounds).  -2: never.
int enqueue_summary(kdop_ctx* c, uint32_t cap, int final_round) {
  __atomic_store_n(&c->hsum->ready, 0u, __ATOMIC_RELEASE);
  hipLaunchKernelGGL(k_sum_counters, dim3(1), dim3(kSumThreads), 0, c->stream, c->dstate, cap, final_round, c->hsum_dev);
  HIPCHK(c, hipGetLastError());
  return KDOP_OK;
}
// Wait for the summary's ready word (k_sum_counters writes it last); the
// stream is polled now and then, so that a failed launch is an error and not
// a hang.  The fold's own clearing of the call state may still run: it is
// ordered before anything queued next on the stream.
int finish_summary(kdop_ctx* c) {
  for (uint64_t spin = 1;; ++spin) {
    if (__atomic_load_n(&c->hsum->ready, __ATOMIC_ACQUIRE)) break;
    if ((spin & 4095u) == 0u) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(c, KDOP_ERR_HIP, "HIP error %d (%s) in the traversal", (int)e, hipGetErrorString(e));
      if (!__atomic_load_n(&c->hsum->ready, __ATOMIC_ACQUIRE)) return fail(c, KDOP_ERR_HIP, "the call summary was not written");
      break;
    }
  }
  c->state_clean = c->hsum->reset != 0;
  return KDOP_OK;
}
// Read n <= 4 device words once the work queued before is done (k_publish_words
// into pinned host memory; the host spins on the sequence number and checks
// the stream now and then, as finish_summary does).
int read_device_words(kdop_ctx* c, const uint32_t* src, uint32_t n, uint32_t* out) {
  const uint32_t seq = ++c->hwords_seq;
  hipLaunchKernelGGL(k_publish_words, dim3(1), dim3(64), 0, c->stream, src, n, c->hwords_dev, seq);
  HIPCHK(c, hipGetLastError());
  for (uint64_t spin = 1;; ++spin) {
    if (__atomic_load_n(&c->hwords->seq, __ATOMIC_ACQUIRE) == seq) break;
    if ((spin & 4095u) == 0u) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(c, KDOP_ERR_HIP, "HIP error %d (%s)", (int)e, hipGetErrorString(e));
      if (__atomic_load_n(&c->hwords->seq, __ATOMIC_ACQUIRE) != seq) return fail(c, KDOP_ERR_HIP, "device words were not published");
      break;
    }
  }
  for (uint32_t i = 0; i < n; ++i) out[i] = c->hwords->v[i];
  return KDOP_OK;
}

int summarize(kdop_ctx* c, uint32_t cap, int final_round) {
  int rc = enqueue_summary(c, cap, final_round);
  return rc ? rc : finish_summary(c);
}

// Sort n keys (a: the keys, b: as many scratch); *result: a or b, whichever
// holds them sorted.
template <class Key>
int device_sort(kdop_ctx* c, Key* a, Key* b, uint32_t n, Key** result) {
  if (n > 1) {
    hipLaunchKernelGGL(k_sort_tiles<Key>, dim3((n + kSortTile - 1) / kSortTile), dim3(kSortThreads), 0, c->stream, a, n);
    for (uint64_t w = kSortTile; w < n; w <<= 1) {
      hipLaunchKernelGGL(k_merge_pass<Key>, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const Key*)a, b, n, (uint32_t)w);
      std::swap(a, b);
    }
  }
  HIPCHK(c, hipGetLastError());
  *result = a;
  return KDOP_OK;
}

// Exclusive scan of n uint32 (sums: ceil(n / kScanBlock) words of scratch;
// total: the sum, on the device).
int device_scan(kdop_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t* total) {
  const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, c->stream, in, out, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, sums, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, c->stream, out, n, (const uint32_t*)sums);
  HIPCHK(c, hipGetLastError());
  return KDOP_OK;
}

// Sort the n contacts by Key (keys: filled by the caller's launch into the
// first half of sortbuf), permute them into contacts_tmp and copy them back.
template <class Key>
int sort_contacts_keys(kdop_ctx* c, uint32_t n) {
  Key* a = reinterpret_cast<Key*>(c->sortbuf.p);
  Key* sorted = nullptr;
  int rc = device_sort(c, a, a + n, n, &sorted);
  if (rc) return rc;
  hipLaunchKernelGGL(k_contact_permute<Key>, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const kdop_contact*)c->contacts.p,
                     (const Key*)sorted, n, c->contacts_tmp.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->contacts.p, c->contacts_tmp.p, (size_t)n * sizeof(kdop_contact), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

// Contacts held in c->contacts: the narrow phase's, or what the reduction left.
uint32_t contact_count(const kdop_ctx* c) {
  return c->reduced ? c->reduced_n : (uint32_t)std::min<size_t>(c->hsum->fills.c, c->contacts.cap);
}

// The scratch laid out from the context's buffers (sized by reduce_scratch).
ReduceScratch reduce_scratch_of(kdop_ctx* c, uint32_t n) {
  ReduceScratch r;
  r.ka = reinterpret_cast<ProxKey*>(c->sortbuf.p);
  r.kb = r.ka + n;
  r.keep = c->redbuf.p;
  r.off = r.keep + n;
  r.sums = r.off + n;
  r.total = r.sums + (n + kScanBlock - 1) / kScanBlock;
  return r;
}

// PROX reduction of n records whose keys fill R.ka: sort the keys, mark the
// points std::unique keeps, and write them to out in sorted order; *kept
// receives their count through pinned host memory once the work is done.
template <class Rec>
int reduce_keys(kdop_ctx* c, const Rec* in, Rec* out, uint32_t n, const ReduceScratch& R, uint32_t* kept) {
  ProxKey* sorted = nullptr;
  int rc = device_sort(c, R.ka, R.kb, n, &sorted);
  if (rc) return rc;
  const dim3 grid((n + 255) / 256);
  hipLaunchKernelGGL(k_prox_unique, grid, dim3(256), 0, c->stream, (const ProxKey*)sorted, n, R.keep);
  HIPCHK(c, hipGetLastError());
  if ((rc = device_scan(c, R.keep, R.off, n, R.sums, R.total))) return rc;
  hipLaunchKernelGGL(k_prox_compact<Rec>, grid, dim3(256), 0, c->stream, in, (const ProxKey*)sorted, (const uint32_t*)R.keep,
                     (const uint32_t*)R.off, n, out);
  HIPCHK(c, hipGetLastError());
  return read_device_words(c, R.total, 1, kept);
}

// Scratch of a reduction of n records in the context's buffers.
int reduce_scratch(kdop_ctx* c, uint32_t n) {
  int rc = ensure(c, c->sortbuf, 2 * (size_t)n * sizeof(ProxKey));
  if (!rc) rc = ensure(c, c->redbuf, 2 * (size_t)n + (n + kScanBlock - 1) / kScanBlock + 1);
  return rc;
}


}  // namespace

extern "C" {

int kdop_traverse(kdop_ctx* c, uint32_t npairs, const uint32_t* body_pairs, uint32_t flags, kdop_counters* out) {
  if (!c) return KDOP_ERR_ARG;
  if (npairs && !body_pairs) return fail(c, KDOP_ERR_ARG, "kdop_traverse: null body_pairs");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  uint32_t levels = 0, level_guess = 0;  // BVTT depth bound; likely depth (deeper tree + 4)
  uint64_t est = 0;
  std::vector<uint2> pairs(npairs);
  uint64_t call_tets = 0;  // tets of both bodies summed over the call's pairs
  for (uint32_t i = 0; i < npairs; ++i) {
    const uint32_t a = body_pairs[2 * i], b = body_pairs[2 * i + 1];
    if (a >= c->bodies.size() || b >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "kdop_traverse: unknown body in pair %u", i);
    pairs[i] = make_uint2(a, b);
    levels = std::max(levels, c->bodies[a].depth + c->bodies[b].depth);
    level_guess = std::max(level_guess, std::max(c->bodies[a].depth, c->bodies[b].depth) + 4);
    est += 4ull * (c->bodies[a].nt + c->bodies[b].nt);
    call_tets += (uint64_t)c->bodies[a].nt + c->bodies[b].nt;
  }
  if (levels > 62) return fail(c, KDOP_ERR_ARG, "kdop_traverse: trees too deep (%u levels)", levels);
  c->have_traversal = c->have_p1 = c->have_narrow = false;
  int rc = KDOP_OK;
  // pair list and seed plan on the device: uploaded again only when they change
  const uint32_t seed_budget = c->seed_budget_set ? c->seed_budget : (npairs >= kManyPairs ? kSeedBudgetManyPairs : kSeedBudgetDefault);
  c->seed_budget_eff = seed_budget;
  const bool same_pairs = c->seed_key.size() == pairs.size() &&
                          (pairs.empty() || std::memcmp(c->seed_key.data(), pairs.data(), pairs.size() * sizeof(uint2)) == 0) &&
                          c->seed_key_budget == seed_budget && c->seed_key_bodies == c->bodies.size();
  uint64_t total_seeds = 0;
  if (same_pairs) {
    total_seeds = c->seed_total;
    for (uint32_t i = 0; i < npairs; ++i) est += c->seed_plan_est[i];
  } else {
    if ((rc = ensure(c, c->dpairs, std::max<uint32_t>(npairs, 1)))) return rc;
    if (npairs) HIPCHK(c, hipMemcpyAsync(c->dpairs.p, pairs.data(), npairs * sizeof(uint2), hipMemcpyHostToDevice, c->stream));
    // seed levels: per body pair the deepest pair of chunk levels whose root
    // pairs fit the per-pair share of the seed budget
    std::vector<SeedDesc> seeds(npairs);
    c->seed_plan_est.assign(npairs, 0);
    const uint64_t share = std::max<uint64_t>(1, seed_budget / std::max<uint32_t>(npairs, 1));
    for (uint32_t i = 0; i < npairs; ++i) {
      const BodyHost& A = c->bodies[pairs[i].x];
      const BodyHost& B = c->bodies[pairs[i].y];
      uint32_t ha = 0, hb = 0;
      uint64_t best = (uint64_t)A.C[0] * B.C[0];
      for (uint32_t x = 0; x < A.H; ++x)
        for (uint32_t y = 0; y < B.H; ++y) {
          const uint64_t prod = (uint64_t)A.C[x] * B.C[y];
          if (prod > share || prod <= best) continue;
          best = prod; ha = x; hb = y;
        }
      seeds[i] = SeedDesc{A.roots_off[ha], A.C[ha], B.roots_off[hb], B.C[hb], (uint32_t)total_seeds, A.node_base, B.node_base,
                          A.C[A.H - 1], B.C[B.H - 1]};
      total_seeds += best;
      c->seed_plan_est[i] = best;
      est += best;
    }
    if (total_seeds >= (1ull << 32)) return fail(c, KDOP_ERR_ARG, "kdop_traverse: too many seed pairs");
    if ((rc = ensure(c, c->dseeds, std::max<uint32_t>(npairs, 1)))) return rc;
    if (npairs) HIPCHK(c, hipMemcpyAsync(c->dseeds.p, seeds.data(), npairs * sizeof(SeedDesc), hipMemcpyHostToDevice, c->stream));
    c->seed_key = pairs;
    c->seed_key_budget = seed_budget;
    c->seed_key_bodies = c->bodies.size();
    c->seed_total = total_seeds;
    ++c->seed_gen;
  }
  const bool keep_p1 = flags & KDOP_KEEP_ALL_PAIRS;
  const bool narrow = flags & KDOP_NARROW;
  // AUTO: the persistent strategy where its depth-first rounds beat one
  // launch per level -- tight DOPs (K >= 12), batches of many pairs at any K,
  // whose heavy pairs the round budget spreads over later rounds, and calls
  // of fewer than kAutoLevelsTets tets, whose BVTT is too narrow to fill the
  // GPU level by level (one 5k-tet pair at K = 8: 0.103 against 0.122 ms per
  // query, 40k tets: 0.164 against 0.199); the level-synchronous one for a
  // few large pairs of loose DOPs (K = 6, 8), whose wide frontiers fill the
  // GPU level by level (two 320k-tet bodies: 0.239-0.248 against 0.250-0.253)
  constexpr uint64_t kAutoLevelsTets = 400000;
  const bool persistent = c->cfg.traversal == KDOP_TRAVERSAL_PERSISTENT ||
                          (c->cfg.traversal == KDOP_TRAVERSAL_AUTO && (c->K >= 12 || npairs >= 64 || call_tets < kAutoLevelsTets));
  est = std::max<uint64_t>(est, 1 << 16);
  if (keep_p1 && (rc = ensure(c, c->p1, est))) return rc;
  if ((rc = ensure(c, c->p2, std::max<uint64_t>(est / 16, 1 << 14)))) return rc;
  if ((rc = ensure(c, c->gate, c->p2.cap))) return rc;
  const uint64_t p2_guess = c->last_p2 ? c->last_p2 : est / 64;
  if (narrow) {
    if ((rc = ensure(c, c->contacts, std::max<uint64_t>(est / 16, 1 << 14)))) return rc;
    if ((rc = ensure(c, c->recs, std::max<uint64_t>(est / 16, 1 << 14)))) return rc;
  }
  if ((rc = ensure(c, c->front[0], est)) || (rc = ensure(c, c->front[1], est))) return rc;
  const uint32_t pwaves = (uint32_t)(c->num_cus * wpc_of(c));
  // the frontier width that ends the ramp: 16 items per persistent wave, 64
  // for calls of many pairs, whose small trees give each pair's subtree to
  // a few blocks unless the frontier is wider (512-pair batch: 16 / 32 / 64
  // / 128 per wave 1.518 / 1.510 / 1.460-1.462 / 1.463-1.466 ms per call; a
  // 64-pair refit series step 0.857-0.860 at 16 against 0.864-0.866 at 64)
  const uint32_t target = c->ramp_target ? c->ramp_target : (npairs >= kManyPairs ? 64u : 16u) * pwaves;
  const uint32_t rounds_async = std::min(std::min(std::max(c->rounds_hint, 1u), kQueuedRounds), c->max_rounds);
  c->rounds_queued = rounds_async;
  uint64_t overflowed = 0;
  uint32_t fallback_levels = 0;  // level-synchronous launches that finished a persistent call
  CallState* d = c->dstate;
  for (int attempt = 0; attempt < 10; ++attempt) {
    if (!c->state_clean) HIPCHK(c, hipMemsetAsync(d, 0, sizeof(CallState), c->stream));
    c->state_clean = false;
    TravArgs T;
    T.nodes = c->nodes.p; T.bodies = c->dbodies.p; T.pairs = c->dpairs.p; T.tops = c->dtops.p;
    T.stripes = d->stripe; T.overflow = &d->overflow;
    T.p1 = keep_p1 ? c->p1.p : nullptr; T.p2 = c->p2.p;
    T.p1_fill = &d->fills.p1; T.p2_fill = &d->fills.p2;
    T.p1_cap = keep_p1 ? (uint32_t)std::min<size_t>(c->p1.cap, 0xFFFFFFFFu) : 0u;
    T.p2_cap = (uint32_t)std::min<size_t>(c->p2.cap, 0xFFFFFFFFu);
    T.npairs = npairs;
    c->round_iters_eff = c->round_iters_set ? c->round_iters
                                            : (npairs >= kRoundItersBatchPairs ? kRoundItersBatch : kRoundItersDefault);
    T.touched = nullptr;
    if (c->count_records && persistent) {
      if ((rc = ensure(c, c->touched, std::max<size_t>(c->nodes_used, 1)))) return rc;
      HIPCHK(c, hipMemsetAsync(c->touched.p, 0, c->nodes_used, c->stream));
      T.touched = c->touched.p;
    }
    const uint32_t cap = (uint32_t)std::min<size_t>(std::min(c->front[0].cap, c->front[1].cap), 0xFFFFFFFFu);
    uint32_t rounds = 0, levels_queued = 0;
    uint32_t last = std::min(levels, c->bfs_max ? c->bfs_max : bfs_auto(call_tets));
    // levels past the one where the last call from the same seeds found its
    // frontier would only find the stop flag (a launch each); round 0 takes
    // whatever level the ramp reaches, so the results do not depend on it
    if (persistent && c->ramp_seed_gen == c->seed_gen && c->ramp_stop_target == target && c->ramp_stop < last)
      last = c->ramp_stop;
    if (npairs) {
      if (persistent) rounds = rounds_async;
      else levels_queued = std::min(levels, level_guess);
    }
    const bool one_pass = narrow_one_pass(c, p2_guess);
    // the persistent ramp's first level goes into the seed kernel when a level
    // is queued at all and the seed pairs are at most a few ramp widths (the
    // tests made are the same either way; when level 0 alone would have ended
    // the ramp, round 0 starts one level deeper)
    const bool seed_expands = persistent && npairs && last >= 1u && total_seeds <= 4ull * target;
    // the launches up to the summary copy: with KDOP_OPT_GRAPHS recorded once
    // into a HIP graph per set of launch arguments and replayed, else (the
    // default, and always with timing events) launched directly; the direct
    // launches now keep up with the GPU (C1 0.093 against 0.094 ms per query
    // with the graph, C2 and the 512-pair batch alike)
    auto enqueue_first = [&]() -> int {
      if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
      if (npairs) {
        dispatch_K<Seed>(c->K, c, T, (uint32_t)total_seeds, &d->level_fill[seed_expands ? 1 : 0], cap, seed_expands);
        if (persistent) {
          // levels 0 (or 1 when the seeds expanded) .. last-1; round 0 takes
          // level `last` when no launch stopped earlier
          for (uint32_t l = seed_expands ? 1u : 0u; l < last; ++l) dispatch_K<Bfs>(c->K, c, T, cap, l, target, last, false);
          if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
          for (uint32_t r = 0; r < rounds_async; ++r) dispatch_K<PersistRound>(c->K, c, T, cap, r, last);
        } else {
          // first the levels the deeper tree suggests; more below if the
          // frontier is not empty by then
          if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
          for (uint32_t l = 0; l <= levels_queued; ++l) dispatch_K<Bfs>(c->K, c, T, cap, l, ~0u, levels, true);
        }
      } else {
        if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
      }
      HIPCHK(c, hipGetLastError());
      if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
      int r2 = KDOP_OK;
      if (narrow && npairs && (r2 = launch_narrow(c, T, persistent ? (int)rounds : -1, !persistent, one_pass))) return r2;
      if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
      return enqueue_summary(c, cap, persistent ? (int)rounds : -1);
    };
    std::vector<uint64_t> sig;
    if (c->use_graphs && c->timing_level == 0 && !c->count_records && npairs) {
      sig = {
          (uint64_t)c->K, npairs, persistent, last, levels_queued, levels, target, cap, total_seeds, narrow, one_pass,
          (uint64_t)(uintptr_t)T.p1, (uint64_t)(uintptr_t)T.p2, T.p1_cap, T.p2_cap, (uint64_t)(uintptr_t)c->nodes.p,
          (uint64_t)(uintptr_t)c->dbodies.p, (uint64_t)(uintptr_t)c->dpairs.p, (uint64_t)(uintptr_t)c->dtops.p,
          (uint64_t)(uintptr_t)c->dseeds.p, (uint64_t)(uintptr_t)c->front[0].p, (uint64_t)(uintptr_t)c->front[1].p,
          (uint64_t)(uintptr_t)c->contacts.p, c->contacts.cap, (uint64_t)(uintptr_t)c->recs.p, c->recs.cap,
          (uint64_t)(uintptr_t)c->gate.p, (uint64_t)wpc_of(c), (uint64_t)ls_of(c), c->keep, (uint64_t)c->num_cus, (uint64_t)bw_of(c), rounds_async,
          (uint64_t)c->item_lanes, c->max_rounds, (uint64_t)c->quad_items, c->round_iters_eff, (uint64_t)c->node_slab};
    }
    const bool replay = !sig.empty() && ((c->tg_exec && sig == c->tg_sig) || sig == c->prev_sig);
    c->prev_sig = sig;
    if (replay) {
      // a capture costs far more than a call: only for launch arguments that repeat
      if (!c->tg_exec || sig != c->tg_sig) {
        if (c->tg_exec) { (void)hipGraphExecDestroy(c->tg_exec); c->tg_exec = nullptr; }
        if (c->tg) { (void)hipGraphDestroy(c->tg); c->tg = nullptr; }
        c->tg_sig.clear();
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rcap = enqueue_first();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (rcap) { if (g) (void)hipGraphDestroy(g); return rcap; }
        if (ec != hipSuccess) return fail(c, KDOP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        c->tg = g;
        HIPCHK(c, hipGraphInstantiate(&c->tg_exec, c->tg, nullptr, nullptr, 0));
        c->tg_sig = sig;
      }
      __atomic_store_n(&c->hsum->ready, 0u, __ATOMIC_RELEASE);
      HIPCHK(c, hipGraphLaunch(c->tg_exec, c->stream));
    } else if ((rc = enqueue_first())) {
      return rc;
    }
    if ((rc = finish_summary(c))) return rc;
    bool grow_front = (c->hsum->overflow & 1u) != 0;
    fallback_levels = 0;
    if (persistent && !grow_front && npairs && c->hsum->round_total[rounds] > 0) {
      // spill left after the queued rounds (the narrow phase above saw it and
      // did nothing): drain it with further rounds, then run the narrow phase.
      // When the rounds run out, the spill becomes a breadth-first level and
      // the level-synchronous launches finish the traversal, however deep.
      bool to_levels = false;
      while (c->hsum->round_total[rounds] > 0 && !(c->hsum->overflow & 1u)) {
        if (rounds >= c->max_rounds) { to_levels = true; break; }
        dispatch_K<PersistRound>(c->K, c, T, cap, rounds, last);
        ++rounds;
        HIPCHK(c, hipGetLastError());
        if ((rc = summarize(c, cap, -2))) return rc;
      }
      grow_front = (c->hsum->overflow & 1u) != 0;
      if (to_levels && !grow_front) {
        // round `rounds` would read buffer (final_level + rounds) & 1; level l
        // reads buffer l & 1, and a level >= 1 reads segments
        const uint32_t l0 = ((c->hsum->final_level + rounds) & 1u) ? 1u : 2u, lend = l0 + levels;
        hipLaunchKernelGGL(k_spill_to_level, dim3(1), dim3(64), 0, c->stream, d, rounds, l0);
        for (uint32_t l = l0; l <= lend; ++l) dispatch_K<Bfs>(c->K, c, T, cap, l, ~0u, lend, true);
        HIPCHK(c, hipGetLastError());
        fallback_levels = lend - l0 + 1;
        if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        if (narrow && (rc = launch_narrow(c, T, -1, true, one_pass))) return rc;
        if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        if ((rc = summarize(c, cap, -1))) return rc;
      } else {
        if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        if (narrow && !grow_front && (rc = launch_narrow(c, T, (int)rounds, false, one_pass))) return rc;
        if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        if ((rc = summarize(c, cap, (int)rounds))) return rc;
      }
      grow_front = (c->hsum->overflow & 1u) != 0;
    }
    if (!persistent && npairs && !grow_front && !c->hsum->final_set) {
      // the level-synchronous frontier outlived the first batch of launches
      for (uint32_t l = levels_queued + 1; l <= levels; ++l) dispatch_K<Bfs>(c->K, c, T, cap, l, ~0u, levels, true);
      HIPCHK(c, hipGetLastError());
      if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
      if (narrow && (rc = launch_narrow(c, T, -1, true, one_pass))) return rc;
      if (c->timing_level >= 1) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
      if ((rc = summarize(c, cap, -1))) return rc;
      grow_front = (c->hsum->overflow & 1u) != 0;
    }
    const CallSummary& h = *c->hsum;
    const Fills& f = h.fills;
    bool again = false;
    if (grow_front) {
      uint64_t mx = 0;
      for (int l = 0; l < 64; ++l) mx = std::max<uint64_t>(mx, h.level_fill[l]);
      for (int r = 1; r < kMaxRounds; ++r) mx = std::max<uint64_t>(mx, (uint64_t)h.round_maxseg[r] * kSegs);
      const size_t want = std::max<size_t>(4 * (size_t)cap, (size_t)mx * 2);
      if ((rc = ensure(c, c->front[0], want)) || (rc = ensure(c, c->front[1], want))) return rc;
      again = true;
    } else {
      if (keep_p1 && f.p1 > T.p1_cap) { if ((rc = ensure(c, c->p1, (size_t)f.p1 + f.p1 / 4))) return rc; again = true; }
      if (f.p2 > T.p2_cap) {
        if ((rc = ensure(c, c->p2, (size_t)f.p2 + f.p2 / 4)) || (rc = ensure(c, c->gate, c->p2.cap))) return rc;
        again = true;
      }
      if (narrow && f.c > c->contacts.cap) { if ((rc = ensure(c, c->contacts, (size_t)f.c + f.c / 4))) return rc; again = true; }
      if (narrow && f.r > c->recs.cap) { if ((rc = ensure(c, c->recs, (size_t)f.r + f.r / 4))) return rc; again = true; }
    }
    if (again) { ++overflowed; continue; }
    float ms[4] = {0, 0, 0, 0};
    if (c->timing_level >= 1) {
      HIPCHK(c, hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]));
      HIPCHK(c, hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]));
      HIPCHK(c, hipEventElapsedTime(&ms[2], c->ev[0], c->ev[2]));
      HIPCHK(c, hipEventElapsedTime(&ms[3], c->ev[3], c->ev[1]));
    }
    c->timing[0] = ms[0]; c->timing[1] = ms[1]; c->timing[2] = ms[2];
    // dominant kernel: k_persist (each round timed by its own dispatch events) or
    // k_bfs_level (levels + 1 launches between two events)
    for (int r = 0; r < kMaxRounds; ++r) c->timing[4 + r] = 0.0;
    if (persistent && c->timing_level >= 2) {
      double sum = 0.0;
      for (uint32_t r = 0; r < rounds; ++r) {
        float t = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&t, c->evr[2 * r], c->evr[2 * r + 1]));
        c->timing[4 + r] = t;
        sum += t;
      }
      c->timing[3] = rounds ? sum / (double)rounds : 0.0;
    } else if (persistent) {
      c->timing[3] = rounds ? ms[3] / (double)rounds : 0.0;  // includes the gaps between rounds (0 without events)
    } else {
      c->timing[3] = ms[3] / (double)(levels + 1);
    }
    c->last_levels = persistent ? 0 : levels;
    c->last_fallback_levels = fallback_levels;
    if (persistent && npairs) {
      c->ramp_seed_gen = c->seed_gen;
      c->ramp_stop_target = target;
      c->ramp_stop = (h.final_set && fallback_levels == 0) ? h.final_level : ~0u;  // (levels after spilled rounds stop elsewhere)
    }
    c->last_distinct_records = 0;
    if (T.touched) {  // count the marked records (a measurement call; the copy is not timed)
      std::vector<uint8_t> h(c->nodes_used);
      HIPCHK(c, hipMemcpy(h.data(), c->touched.p, h.size(), hipMemcpyDeviceToHost));
      for (const uint8_t v : h) c->last_distinct_records += v ? 1u : 0u;
    }
    c->last_rounds = persistent ? rounds : 0;
    c->last_persistent = persistent;
    c->last_npairs = npairs;
    c->last_p2 = f.p2;
    if (persistent && npairs) {  // rounds this query needed: queue as many next time
      uint32_t needed = 1;
      for (uint32_t r = 1; r < (uint32_t)kMaxRounds; ++r)
        if (h.round_total[r] > 0) needed = r + 1;
      needed = std::min(needed, kQueuedRounds);
      if (needed >= c->rounds_hint) {
        c->rounds_hint = needed;
        c->rounds_calm = 0;
      } else if (++c->rounds_calm >= 8u) {
        c->rounds_hint -= 1u;
        c->rounds_calm = 0;
      }
    }
    c->last_pairs = pairs;
    c->have_traversal = true;
    c->have_p1 = keep_p1;
    c->have_narrow = narrow;
    c->reduced = false;
    if (out) {
      const unsigned long long* t = h.total;  // CounterStripe field order
      std::memset(out, 0, sizeof(*out));
      out->tests = t[0];
      out->descent_tests = t[1];
      out->branch_overlaps = t[2];
      out->leaf_pairs = t[3];
      out->surface_pairs = f.p2;
      out->contacts = narrow ? f.c : 0;
      out->narrow_pairs = narrow ? t[4] : 0;
      out->branch_pairs = t[5];
      out->reference_tests = t[6] + t[5] + t[1];
      out->overflow = overflowed;
    }
    return KDOP_OK;
  }
  return fail(c, KDOP_ERR_CAPACITY, "kdop_traverse: buffers kept overflowing");
}

int kdop_get_pairs(kdop_ctx* c, int which, uint32_t* dst, uint64_t cap, int dst_is_device, uint64_t* n) {
  if (!c || !n) return KDOP_ERR_ARG;
  if (!c->have_traversal) return fail(c, KDOP_ERR_STATE, "kdop_get_pairs: no traversal yet");
  if (which == 0 && !c->have_p1) return fail(c, KDOP_ERR_STATE, "kdop_get_pairs: P1 not kept (pass KDOP_KEEP_ALL_PAIRS)");
  const Fills& f = c->hsum->fills;
  const uint64_t cnt = which == 0 ? f.p1 : f.p2;
  *n = cnt;
  if (!dst) return KDOP_OK;
  if (cap < cnt) return fail(c, KDOP_ERR_CAPACITY, "kdop_get_pairs: need %llu pairs", (unsigned long long)cnt);
  if (cnt == 0) return KDOP_OK;
  const LeafPair* src = which == 0 ? c->p1.p : c->p2.p;
  HIPCHK(c, hipMemcpyAsync(dst, src, cnt * sizeof(LeafPair), dst_is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

int kdop_get_contacts(kdop_ctx* c, kdop_contact* dst, uint64_t cap, int dst_is_device, uint64_t* n) {
  if (!c || !n) return KDOP_ERR_ARG;
  if (!c->have_narrow) return fail(c, KDOP_ERR_STATE, "kdop_get_contacts: last traversal ran without KDOP_NARROW");
  const uint32_t nc = c->reduced ? c->reduced_n : c->hsum->fills.c;
  *n = nc;
  if (!dst) return KDOP_OK;
  if (cap < nc) return fail(c, KDOP_ERR_CAPACITY, "kdop_get_contacts: need %u contacts", nc);
  if (nc == 0) return KDOP_OK;
  HIPCHK(c, hipMemcpyAsync(dst, c->contacts.p, (size_t)nc * sizeof(kdop_contact), dst_is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

int kdop_get_world(kdop_ctx* c, uint32_t body, float* X, float* Y, float* Z) {
  if (!c || !X || !Y || !Z) return KDOP_ERR_ARG;
  if (body >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "unknown body %u", body);
  const BodyHost& b = c->bodies[body];
  HIPCHK(c, hipMemcpyAsync(X, b.X, b.nv * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Y, b.Y, b.nv * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Z, b.Z, b.nv * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

int kdop_export_tree(kdop_ctx* c, uint32_t body, uint64_t cap, float* vol, uint32_t* range, uint8_t* kind, float* root, uint64_t* n) {
  if (!c || !n) return KDOP_ERR_ARG;
  if (body >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "unknown body %u", body);
  const BodyHost& b = c->bodies[body];
  *n = b.real_nodes;
  if (!vol && !range && !kind && !root) return KDOP_OK;
  if (cap < *n) return fail(c, KDOP_ERR_CAPACITY, "kdop_export_tree: need %llu nodes", (unsigned long long)b.real_nodes);
  const int K = c->K, S = rec_stride(K);
  std::vector<float> rec((size_t)b.nnodes * S);
  HIPCHK(c, hipMemcpyAsync(rec.data(), c->nodes.p + (size_t)b.node_base * S, rec.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto meta = [&](uint32_t local, int w) {
    uint32_t u;
    std::memcpy(&u, &rec[(size_t)local * S + K + w], 4);
    return u;
  };
  // tet ranges: children always sit in later slots (within a chunk in BFS
  // order, across chunks on lower levels), so a reverse sweep sees them first
  std::vector<uint32_t> lo(b.nnodes, 0), hi(b.nnodes, 0);
  for (int64_t i = (int64_t)b.nnodes - 1; i >= 1; --i) {
    const uint32_t m0 = meta((uint32_t)i, 0), m1 = meta((uint32_t)i, 1);
    if (m0 == 0xFFFFFFFFu) continue;  // free slot
    if (m0 & F_LEAF) { lo[i] = hi[i] = m0 & IDX_MASK; continue; }
    const uint32_t c0 = (m0 & IDX_MASK) - b.node_base, c1 = (m1 & IDX_MASK) - b.node_base;
    lo[i] = lo[c0]; hi[i] = hi[c1];
  }
  uint64_t k_out = 0;
  for (uint32_t i = 1; i < b.nnodes; ++i) {
    const uint32_t m0 = meta(i, 0);
    if (m0 == 0xFFFFFFFFu) continue;
    if (vol) for (int k = 0; k < K; ++k) vol[(size_t)k_out * K + k] = rec[(size_t)i * S + k];
    if (range) { range[2 * k_out] = lo[i]; range[2 * k_out + 1] = hi[i]; }
    if (kind) kind[k_out] = (m0 & F_LEAF) ? 2 : ((m0 & F_BRANCH) ? 1 : 0);
    ++k_out;
  }
  if (root) for (int k = 0; k < K; ++k) root[k] = rec[k];
  return KDOP_OK;
}

int kdop_traversal_stats(kdop_ctx* c, uint64_t* out) {
  if (!c || !out) return KDOP_ERR_ARG;
  std::memset(out, 0, 20 * sizeof(uint64_t));
  out[10] = (uint64_t)ls_of(c);
  out[11] = (uint64_t)wpc_of(c);
  out[13] = (uint64_t)c->item_lanes;
  out[14] = (uint64_t)bw_of(c);
  if (!c->have_traversal || !c->last_persistent) return KDOP_OK;
  const CallSummary& S = *c->hsum;
  out[0] = S.final_level; out[1] = S.final_count; out[2] = c->last_rounds; out[3] = S.total[7];
  out[4] = S.total[8]; out[5] = S.total[13]; out[6] = S.total[9]; out[7] = (uint64_t)c->num_cus * wpc_of(c);
  out[8] = S.total[10]; out[9] = c->last_rounds;
  out[12] = S.total[11];
  out[15] = c->last_fallback_levels;
  out[16] = S.total[12];
  out[17] = c->last_distinct_records;
  return KDOP_OK;
}

int kdop_last_timing(kdop_ctx* c, double* out) {
  if (!c || !out) return KDOP_ERR_ARG;
  for (int i = 0; i < 4 + kMaxRounds; ++i) out[i] = c->timing[i];
  return KDOP_OK;
}

int kdop_gather_roof(kdop_ctx* c, uint64_t window, uint32_t nloads, uint32_t seed, double* gbytes_per_s) {
  if (!c || !gbytes_per_s) return KDOP_ERR_ARG;
  *gbytes_per_s = 0.0;
  if (c->nodes_used < 2) return fail(c, KDOP_ERR_STATE, "kdop_gather_roof: no bodies");
  if (nloads < 1024 || nloads > (1u << 26)) return fail(c, KDOP_ERR_ARG, "kdop_gather_roof: nloads %u not in [1024, 2^26]", nloads);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint64_t span = std::min<uint64_t>(window ? window : c->nodes_used, c->nodes_used);
  std::vector<uint32_t> h(nloads);
  uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
  for (uint32_t i = 0; i < nloads; ++i) {  // xorshift64*, uniform slots of the window
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
    h[i] = (uint32_t)(((x * 0x2545F4914F6CDD1Dull) >> 32) % span);
  }
  const uint32_t blocks = (uint32_t)c->num_cus * 16u;
  uint32_t *didx = nullptr, *dout = nullptr;
  if (hipMalloc(&didx, 4 * (size_t)nloads) != hipSuccess) return fail(c, KDOP_ERR_NOMEM, "kdop_gather_roof: hipMalloc");
  if (hipMalloc(&dout, 4 * (size_t)blocks * 256) != hipSuccess) { (void)hipFree(didx); return fail(c, KDOP_ERR_NOMEM, "kdop_gather_roof: hipMalloc"); }
  int rc = KDOP_OK;
  std::vector<float> ms;
  if (hipMemcpy(didx, h.data(), 4 * (size_t)nloads, hipMemcpyHostToDevice) != hipSuccess) rc = fail(c, KDOP_ERR_HIP, "kdop_gather_roof: copy");
  for (int rep = 0; rep < 7 && rc == KDOP_OK; ++rep) {  // one warm-up, then the median of six
    if (hipEventRecord(c->ev[0], c->stream) != hipSuccess) { rc = fail(c, KDOP_ERR_HIP, "kdop_gather_roof: event"); break; }
    dispatch_K<GatherProbe>(c->K, c, (const uint32_t*)didx, nloads, dout, blocks);
    if (hipEventRecord(c->ev[1], c->stream) != hipSuccess || hipEventSynchronize(c->ev[1]) != hipSuccess) {
      rc = fail(c, KDOP_ERR_HIP, "kdop_gather_roof: %s", hipGetErrorString(hipGetLastError()));
      break;
    }
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
    if (rep) ms.push_back(t);
  }
  (void)hipFree(didx);
  (void)hipFree(dout);
  if (rc) return rc;
  std::sort(ms.begin(), ms.end());
  const double med = 0.5 * (ms[2] + ms[3]);
  const double rec_bytes = 16.0 * ((c->K + 2 + 3) / 4);  // the words a record's load reads: K volumes and 2 meta words
  *gbytes_per_s = (double)(nloads / 4 * 4) * rec_bytes / (med * 1e-3) / 1e9;
  return KDOP_OK;
}

int kdop_narrow_adjoint(kdop_ctx* c, uint32_t nbodies, const uint32_t* bodies, const float* dphi_dev, float* phi_dev,
                        float* grad_pos_dev, float* grad_quat_dev) {
  if (!c) return KDOP_ERR_ARG;
  if (!c->have_narrow) return fail(c, KDOP_ERR_STATE, "kdop_narrow_adjoint: needs a kdop_traverse with KDOP_NARROW first");
  if (nbodies && (!bodies || !grad_pos_dev || !grad_quat_dev)) return fail(c, KDOP_ERR_ARG, "kdop_narrow_adjoint: null argument");
  for (uint32_t i = 0; i < nbodies; ++i)
    if (bodies[i] >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "kdop_narrow_adjoint: unknown body %u", bodies[i]);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t npairs = c->last_npairs;
  int rc = ensure(c, c->adj_acc, std::max<size_t>((size_t)npairs * kAdjStripes * kAdjVals, 1));
  if (!rc) rc = ensure(c, c->adj_bodies, std::max<uint32_t>(nbodies, 1));
  if (rc) return rc;
  HIPCHK(c, hipMemsetAsync(c->adj_acc.p, 0, (size_t)npairs * kAdjStripes * kAdjVals * sizeof(double), c->stream));
  if (nbodies) HIPCHK(c, hipMemcpyAsync(c->adj_bodies.p, bodies, nbodies * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  AdjArgs A;
  A.bodies = c->dbodies.p; A.pairs = c->dpairs.p; A.recs = c->recs.p;
  A.nrec = (uint32_t)std::min<size_t>(c->hsum->fills.r, c->recs.cap);
  A.poses = c->dposes.p; A.dphi = dphi_dev; A.acc = c->adj_acc.p;
  if (npairs && c->hsum->fills.r) hipLaunchKernelGGL(k_adjoint, dim3((unsigned)(c->num_cus * 4)), dim3(256), 0, c->stream, A);
  const uint32_t nthreads = std::max(7 * nbodies, npairs);
  if (nthreads)
    hipLaunchKernelGGL(k_adjoint_finish, dim3((nthreads + 255) / 256), dim3(256), 0, c->stream, c->dpairs.p, npairs, c->adj_acc.p,
                       c->adj_bodies.p, nbodies, grad_pos_dev, grad_quat_dev, phi_dev);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

int kdop_sort_contacts(kdop_ctx* c, int order) {
  if (!c) return KDOP_ERR_ARG;
  if (!c->have_narrow) return fail(c, KDOP_ERR_STATE, "kdop_sort_contacts: last traversal ran without KDOP_NARROW");
  if (order != KDOP_ORDER_SEQUENTIAL && order != KDOP_ORDER_OPENCL) return fail(c, KDOP_ERR_ARG, "kdop_sort_contacts: unknown order %d", order);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t n = contact_count(c);
  if (n < 2) return KDOP_OK;
  int rc = ensure(c, c->sortbuf, 2 * (size_t)n * std::max(sizeof(ContactSeqKey), sizeof(ContactClKey)));
  if (!rc) rc = ensure(c, c->contacts_tmp, n);
  if (rc) return rc;
  const dim3 grid((n + 255) / 256);
  if (order == KDOP_ORDER_SEQUENTIAL) {
    hipLaunchKernelGGL(k_contact_keys_seq, grid, dim3(256), 0, c->stream, (const kdop_contact*)c->contacts.p, n, (const uint2*)c->dpairs.p,
                       (const BodyDev*)c->dbodies.p, reinterpret_cast<ContactSeqKey*>(c->sortbuf.p));
    return sort_contacts_keys<ContactSeqKey>(c, n);
  }
  hipLaunchKernelGGL(k_contact_keys_cl, grid, dim3(256), 0, c->stream, (const kdop_contact*)c->contacts.p, n,
                     reinterpret_cast<ContactClKey*>(c->sortbuf.p));
  return sort_contacts_keys<ContactClKey>(c, n);
}

int kdop_reduce_contacts(kdop_ctx* c, uint64_t* n) {
  if (!c || !n) return KDOP_ERR_ARG;
  if (!c->have_narrow) return fail(c, KDOP_ERR_STATE, "kdop_reduce_contacts: last traversal ran without KDOP_NARROW");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t nc = contact_count(c);
  if (nc == 0) {
    c->reduced_n = 0;
    c->reduced = true;
    *n = 0;
    return KDOP_OK;
  }
  int rc = reduce_scratch(c, nc);
  if (!rc) rc = ensure(c, c->contacts_tmp, nc);
  if (rc) return rc;
  hipLaunchKernelGGL(k_prox_keys_pairs, dim3((nc + 255) / 256), dim3(256), 0, c->stream, (const kdop_contact*)c->contacts.p, nc,
                     (const uint2*)c->dpairs.p, reinterpret_cast<ProxKey*>(c->sortbuf.p));
  uint32_t kept = 0;
  if ((rc = reduce_keys<kdop_contact>(c, c->contacts.p, c->contacts_tmp.p, nc, reduce_scratch_of(c, nc), &kept))) return rc;
  if (kept) HIPCHK(c, hipMemcpyAsync(c->contacts.p, c->contacts_tmp.p, (size_t)kept * sizeof(kdop_contact), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->reduced_n = kept;
  c->reduced = true;
  *n = kept;
  return KDOP_OK;
}

int kdop_reduce_prox_contacts(kdop_ctx* c, uint64_t n, const kdop_prox_contact* in, kdop_prox_contact* out, int is_device, uint64_t* n_out) {
  if (!c || !n_out) return KDOP_ERR_ARG;
  *n_out = 0;
  if (n == 0) return KDOP_OK;
  if (!in || !out) return fail(c, KDOP_ERR_ARG, "kdop_reduce_prox_contacts: null argument");
  if (n >= 0xFFFFFFFFull) return fail(c, KDOP_ERR_ARG, "kdop_reduce_prox_contacts: %llu contacts (at most 2^32 - 2)", (unsigned long long)n);
  if (in == out) return fail(c, KDOP_ERR_ARG, "kdop_reduce_prox_contacts: out aliases in");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t nc = (uint32_t)n;
  int rc = reduce_scratch(c, nc);
  if (!rc && !is_device) rc = ensure(c, c->prox_io, 2 * (size_t)nc);
  if (rc) return rc;
  const kdop_prox_contact* din = in;
  kdop_prox_contact* dout = out;
  if (!is_device) {
    HIPCHK(c, hipMemcpyAsync(c->prox_io.p, in, (size_t)nc * sizeof(kdop_prox_contact), hipMemcpyHostToDevice, c->stream));
    din = c->prox_io.p;
    dout = c->prox_io.p + nc;
  }
  hipLaunchKernelGGL(k_prox_keys_points, dim3((nc + 255) / 256), dim3(256), 0, c->stream, din, nc, reinterpret_cast<ProxKey*>(c->sortbuf.p));
  uint32_t kept = 0;
  if ((rc = reduce_keys<kdop_prox_contact>(c, din, dout, nc, reduce_scratch_of(c, nc), &kept))) return rc;
  if (!is_device && kept) {
    HIPCHK(c, hipMemcpyAsync(out, dout, (size_t)kept * sizeof(kdop_prox_contact), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  *n_out = kept;
  return KDOP_OK;
}

int kdop_reduce_prox_contacts_noalloc(kdop_ctx* c, uint64_t n, const kdop_prox_contact* in_dev, kdop_prox_contact* out_dev,
                                      uint64_t* n_out) {
  if (!c || !n_out) return KDOP_ERR_ARG;
  *n_out = 0;
  if (n == 0) return KDOP_OK;
  if (!in_dev || !out_dev) return fail(c, KDOP_ERR_ARG, "kdop_reduce_prox_contacts_noalloc: null argument");
  if (n > KDOP_NOALLOC_MAX_CONTACTS)
    return fail(c, KDOP_ERR_CAPACITY, "kdop_reduce_prox_contacts_noalloc: %llu contacts, at most %u (KDOP_NOALLOC_MAX_CONTACTS)",
                (unsigned long long)n, (unsigned)KDOP_NOALLOC_MAX_CONTACTS);
  if (in_dev == out_dev) return fail(c, KDOP_ERR_ARG, "kdop_reduce_prox_contacts_noalloc: out aliases in");
  if (!c->scratch) return fail(c, KDOP_ERR_HIP, "kdop_reduce_prox_contacts_noalloc: no scratch on this device (%s)", c->scratch_err.c_str());
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t nc = (uint32_t)n;
  const ReduceScratch R = noalloc_scratch_of(c->scratch);
  hipLaunchKernelGGL(k_prox_keys_points, dim3((nc + 255) / 256), dim3(256), 0, c->stream, in_dev, nc, R.ka);
  uint32_t kept = 0;
  const int rc = reduce_keys<kdop_prox_contact>(c, in_dev, out_dev, nc, R, &kept);
  if (rc) return rc;
  *n_out = kept;
  return KDOP_OK;
}

int kdop_mesh_reorder_gpu(kdop_ctx* c, uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt, const uint32_t* tets,
                          float* Xo, float* Yo, float* Zo, uint32_t* tets_o) {
  if (!c) return KDOP_ERR_ARG;
  if (!X || !Y || !Z || !tets || !Xo || !Yo || !Zo || !tets_o) return fail(c, KDOP_ERR_ARG, "kdop_mesh_reorder_gpu: null argument");
  if (nv == 0 || nt == 0) return fail(c, KDOP_ERR_ARG, "kdop_mesh_reorder_gpu: empty mesh (nv=%u nt=%u)", nv, nt);
  if (nt >= (1u << 29) || nv >= (1u << 31)) return fail(c, KDOP_ERR_ARG, "kdop_mesh_reorder_gpu: mesh too large (nv=%u nt=%u)", nv, nt);
  for (size_t i = 0; i < 4 * (size_t)nt; ++i)
    if (tets[i] >= nv) return fail(c, KDOP_ERR_ARG, "kdop_mesh_reorder_gpu: tet vertex index %u out of range (nv=%u)", tets[i], nv);
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const uint32_t ns = 4 * nt, nbt = (nt + 255) / 256;
  const uint32_t nsum = (std::max(ns, nv) + kScanBlock - 1) / kScanBlock;
  // one allocation, carved into 256-byte aligned pieces
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  const size_t oX = take(3 * (size_t)nv * 4), oXo = take(3 * (size_t)nv * 4), oT = take((size_t)ns * 4), oTo = take((size_t)ns * 4);
  const size_t oK = take(2 * (size_t)nt * sizeof(MortonKey)), oPart = take((size_t)nbt * 6 * 4), oP = take(sizeof(ReorderParams));
  const size_t oFirst = take((size_t)nv * 4), oFlag = take((size_t)ns * 4), oFlagS = take((size_t)ns * 4);
  const size_t oUnused = take((size_t)nv * 4), oUnusedS = take((size_t)nv * 4), oSums = take((size_t)nsum * 4), oTot = take(8);
  unsigned char* m = nullptr;
  {
    const hipError_t e = hipMalloc(&m, off);
    if (e != hipSuccess) return fail(c, KDOP_ERR_NOMEM, "kdop_mesh_reorder_gpu: hipMalloc(%zu bytes) failed", off);
  }
  float* dX = reinterpret_cast<float*>(m + oX);
  float* dY = dX + nv;
  float* dZ = dY + nv;
  float* dXo = reinterpret_cast<float*>(m + oXo);
  float* dYo = dXo + nv;
  float* dZo = dYo + nv;
  uint32_t* dT = reinterpret_cast<uint32_t*>(m + oT);
  uint32_t* dTo = reinterpret_cast<uint32_t*>(m + oTo);
  MortonKey* ka = reinterpret_cast<MortonKey*>(m + oK);
  MortonKey* kb = ka + nt;
  float* part = reinterpret_cast<float*>(m + oPart);
  ReorderParams* P = reinterpret_cast<ReorderParams*>(m + oP);
  uint32_t* first = reinterpret_cast<uint32_t*>(m + oFirst);
  uint32_t* flag = reinterpret_cast<uint32_t*>(m + oFlag);
  uint32_t* flag_s = reinterpret_cast<uint32_t*>(m + oFlagS);
  uint32_t* unused = reinterpret_cast<uint32_t*>(m + oUnused);
  uint32_t* unused_s = reinterpret_cast<uint32_t*>(m + oUnusedS);
  uint32_t* sums = reinterpret_cast<uint32_t*>(m + oSums);
  uint32_t* tot = reinterpret_cast<uint32_t*>(m + oTot);
  int rc = KDOP_OK;
  auto run = [&]() -> int {
    HIPCHK(c, hipMemcpyAsync(dX, X, (size_t)nv * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dY, Y, (size_t)nv * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dZ, Z, (size_t)nv * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dT, tets, (size_t)ns * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(first, 0xFF, (size_t)nv * 4, c->stream));
    hipLaunchKernelGGL(k_reorder_bounds, dim3(nbt), dim3(256), 0, c->stream, (const float*)dX, (const float*)dY, (const float*)dZ,
                       (const uint32_t*)dT, nt, part);
    hipLaunchKernelGGL(k_reorder_scale, dim3(1), dim3(256), 0, c->stream, (const float*)part, nbt, P);
    hipLaunchKernelGGL(k_reorder_codes, dim3(nbt), dim3(256), 0, c->stream, (const float*)dX, (const float*)dY, (const float*)dZ,
                       (const uint32_t*)dT, nt, (const ReorderParams*)P, ka);
    HIPCHK(c, hipGetLastError());
    MortonKey* sorted = nullptr;
    int r = device_sort(c, ka, kb, nt, &sorted);
    if (r) return r;
    const uint32_t nmax = std::max(ns, nv);
    hipLaunchKernelGGL(k_reorder_first, dim3((ns + 255) / 256), dim3(256), 0, c->stream, (const MortonKey*)sorted, (const uint32_t*)dT, nt,
                       first);
    hipLaunchKernelGGL(k_reorder_flags, dim3((nmax + 255) / 256), dim3(256), 0, c->stream, (const MortonKey*)sorted, (const uint32_t*)dT, nt,
                       (const uint32_t*)first, nv, flag, unused);
    HIPCHK(c, hipGetLastError());
    if ((r = device_scan(c, flag, flag_s, ns, sums, tot))) return r;
    if ((r = device_scan(c, unused, unused_s, nv, sums, tot + 1))) return r;
    hipLaunchKernelGGL(k_reorder_tets, dim3((ns + 255) / 256), dim3(256), 0, c->stream, (const MortonKey*)sorted, (const uint32_t*)dT, nt,
                       (const uint32_t*)first, (const uint32_t*)flag_s, dTo);
    hipLaunchKernelGGL(k_reorder_verts, dim3((nv + 255) / 256), dim3(256), 0, c->stream, (const float*)dX, (const float*)dY, (const float*)dZ,
                       nv, (const uint32_t*)first, (const uint32_t*)flag_s, (const uint32_t*)unused_s, (const uint32_t*)tot, dXo, dYo, dZo);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(Xo, dXo, (size_t)nv * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Yo, dYo, (size_t)nv * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Zo, dZo, (size_t)nv * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(tets_o, dTo, (size_t)ns * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return KDOP_OK;
  };
  rc = run();
  if (rc) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(m);
  return rc;
}

int kdop_raycast(kdop_ctx* c, uint32_t body, uint32_t n, const float* origins, const float* directions, float* hit_points, float* lengths,
                 uint8_t* hits, int is_device) {
  if (!c) return KDOP_ERR_ARG;
  if (body >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "kdop_raycast: unknown body %u", body);
  if (n == 0) return KDOP_OK;
  if (!origins || !directions || !hit_points || !lengths || !hits) return fail(c, KDOP_ERR_ARG, "kdop_raycast: null argument");
  if (n >= (1u << 28)) return fail(c, KDOP_ERR_ARG, "kdop_raycast: too many rays (%u)", n);
  const BodyHost& b = c->bodies[body];
  uint32_t lg = 0;
  while ((1u << lg) < b.L) ++lg;
  if (lg + 1u > kChunkStack) return fail(c, KDOP_ERR_ARG, "kdop_raycast: chunk trees too deep (%u leaves)", b.L);
  // below the top-level roots every item is a branch-chunk root after (H - 1)
  // log2 L levels, taken two per step
  const uint32_t steps = ((b.H - 1u) * std::max(lg, 1u) + 1u) / 2u;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  if (is_device && ((reinterpret_cast<uintptr_t>(origins) | reinterpret_cast<uintptr_t>(directions) | reinterpret_cast<uintptr_t>(hit_points) |
                     reinterpret_cast<uintptr_t>(lengths)) & 3u))
    return fail(c, KDOP_ERR_ARG, "kdop_raycast: device float arrays must be 4-byte aligned");
  // device arrays are read and written in place; host arrays go through raybuf
  const float* d_o = origins;
  const float* d_d = directions;
  float* d_h = hit_points;
  float* d_l = lengths;
  uint8_t* d_f = hits;
  int rc = KDOP_OK;
  if (!is_device) {
    const size_t bytes_in = 6 * (size_t)n * sizeof(float), bytes_out = 4 * (size_t)n * sizeof(float);
    if ((rc = ensure(c, c->raybuf, bytes_in + bytes_out + n))) return rc;
    float* s_o = reinterpret_cast<float*>(c->raybuf.p);
    float* s_d = s_o + 3 * (size_t)n;
    d_h = s_d + 3 * (size_t)n;
    d_l = d_h + 3 * (size_t)n;
    d_f = reinterpret_cast<uint8_t*>(d_l + n);
    HIPCHK(c, hipMemcpyAsync(s_o, origins, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s_d, directions, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    d_o = s_o;
    d_d = s_d;
  }
  // scan scratch for m values: [count m][offset m][block sums][total][bad]
  auto scratch = [&](uint32_t m) { return ensure(c, c->sg_scan, 2 * (size_t)m + (m + kScanBlock - 1) / kScanBlock + 2); };
  auto read_words = [&](uint32_t m, uint32_t* host, uint32_t words) -> int {
    const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
    return read_device_words(c, c->sg_scan.p + 2 * (size_t)m + nb, words, host);
  };
  auto scan = [&](uint32_t m) {
    const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
    uint32_t* cnt = c->sg_scan.p;
    return device_scan(c, cnt, cnt + m, m, cnt + 2 * (size_t)m, cnt + 2 * (size_t)m + nb);
  };
  // rays go in batches whose (ray, node) frontier stays within the item limit
  const uint32_t max_items = c->sg_max_items;
  uint32_t s0 = 0, batch = n;
  while (s0 < n) {
    const uint32_t sn = std::min(batch, n - s0);
    const float* bo = d_o + 3 * (size_t)s0;
    const float* bd = d_d + 3 * (size_t)s0;
    if ((rc = scratch(sn))) return rc;
    dispatch_K<RcSeedCount>(c->K, c, body, sn, bo, bd, c->sg_scan.p);
    HIPCHK(c, hipGetLastError());
    if ((rc = scan(sn))) return rc;
    uint32_t m = 0;
    if ((rc = read_words(sn, &m, 1))) return rc;
    bool too_big = m > max_items;
    if (m && !too_big) {
      if ((rc = ensure(c, c->sg_items[0], m))) return rc;
      hipLaunchKernelGGL(k_sg_seed_write, dim3((sn + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, c->dtops.p, sn,
                         (const uint32_t*)c->sg_scan.p, (const uint32_t*)(c->sg_scan.p + sn), c->sg_items[0].p);
      HIPCHK(c, hipGetLastError());
    }
    int cur = 0;
    for (uint32_t step = 0; !too_big && m && step < steps; ++step) {
      if ((rc = scratch(m))) return rc;
      if ((rc = ensure(c, c->sg_kind, m))) return rc;
      dispatch_K<RcExpand2Count>(c->K, c, (const uint2*)c->sg_items[cur].p, m, bo, bd, c->sg_scan.p, c->sg_kind.p);
      HIPCHK(c, hipGetLastError());
      if ((rc = scan(m))) return rc;
      uint32_t m2 = 0;
      if ((rc = read_words(m, &m2, 1))) return rc;
      if (m2 > max_items) { too_big = true; break; }
      if (m2) {
        if ((rc = ensure(c, c->sg_items[cur ^ 1], m2))) return rc;
        dispatch_K<SgExpand2Write>(c->K, c, (const uint2*)c->sg_items[cur].p, m, (const uint8_t*)c->sg_kind.p,
                                   (const uint32_t*)(c->sg_scan.p + m), c->sg_items[cur ^ 1].p);
        HIPCHK(c, hipGetLastError());
      }
      cur ^= 1;
      m = m2;
    }
    if (too_big) {
      if (sn == 1) return fail(c, KDOP_ERR_CAPACITY, "kdop_raycast: ray %u meets more than %u nodes of one level", s0, max_items);
      batch = std::max(1u, sn / 2);
      continue;
    }
    // per item: its chunk's best length; per ray: its first item; the bad
    // flag; the passing roots' item list (the other frontier buffer)
    if ((rc = scratch(m)) || (rc = ensure(c, c->rc_best, std::max<size_t>(m, 1))) || (rc = ensure(c, c->rc_first, (size_t)sn + 1)))
      return rc;
    float* best = c->rc_best.p;
    uint32_t* first = c->rc_first.p;
    uint32_t* bad = first + sn;
    HIPCHK(c, hipMemsetAsync(first, 0xFF, 4 * (size_t)sn, c->stream));
    HIPCHK(c, hipMemsetAsync(bad, 0, 4, c->stream));
    if (m) {
      const uint2* items = (const uint2*)c->sg_items[cur].p;
      dispatch_K<RcRoots>(c->K, c, items, m, bo, bd, c->sg_scan.p, best, bad);
      HIPCHK(c, hipGetLastError());
      if ((rc = scan(m))) return rc;
      uint32_t np = 0;
      if ((rc = read_words(m, &np, 1))) return rc;
      if (np) {
        if ((rc = ensure(c, c->sg_items[cur ^ 1], (np + 1) / 2))) return rc;
        uint32_t* list = reinterpret_cast<uint32_t*>(c->sg_items[cur ^ 1].p);
        hipLaunchKernelGGL(k_gather_flagged, dim3((m + 255) / 256), dim3(256), 0, c->stream, (const uint32_t*)c->sg_scan.p,
                           (const uint32_t*)(c->sg_scan.p + m), m, list);
        dispatch_K<RcChunks>(c->K, c, body, items, (const uint32_t*)list, np, bo, bd, best);
      }
      hipLaunchKernelGGL(k_rc_first, dim3((m + 255) / 256), dim3(256), 0, c->stream, items, m, first);
      HIPCHK(c, hipGetLastError());
    }
    dispatch_K<RcResolve>(c->K, c, body, (const uint2*)c->sg_items[cur].p, m, (const float*)best, (const uint32_t*)first, sn, bo, bd, s0,
                          d_h, d_l, d_f);
    HIPCHK(c, hipGetLastError());
    uint32_t bad_h = 0;
    if ((rc = read_device_words(c, bad, 1, &bad_h))) return rc;
    if (bad_h) return fail(c, KDOP_ERR_INTERNAL, "kdop_raycast: a node other than a branch-chunk root after %u steps", steps);
    s0 += sn;
  }
  if (!is_device) {
    HIPCHK(c, hipMemcpyAsync(hit_points, d_h, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(lengths, d_l, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hits, d_f, n, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

int kdop_single_traversal(kdop_ctx* c, uint32_t body, uint32_t n, const float* spheres, int flip, kdop_sphere_contact* out, uint64_t cap,
                          uint64_t* count, int is_device) {
  if (!c || !count) return KDOP_ERR_ARG;
  *count = 0;
  if (body >= c->bodies.size()) return fail(c, KDOP_ERR_ARG, "kdop_single_traversal: unknown body %u", body);
  if (n == 0) return KDOP_OK;
  if (!spheres) return fail(c, KDOP_ERR_ARG, "kdop_single_traversal: null spheres");
  if (n >= (1u << 28)) return fail(c, KDOP_ERR_ARG, "kdop_single_traversal: too many spheres (%u)", n);
  const BodyHost& b = c->bodies[body];
  uint32_t lg = 0;
  while ((1u << lg) < b.L) ++lg;
  // below the top-level chunk roots every item is a leaf after H log2 L levels,
  // taken two per step
  const uint32_t steps = (b.H * std::max(lg, 1u) + 1u) / 2u;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  int rc = KDOP_OK;
  if (is_device && (reinterpret_cast<uintptr_t>(spheres) & 15u))
    return fail(c, KDOP_ERR_ARG, "kdop_single_traversal: device spheres must be 16-byte aligned");
  const float* d_s = spheres;  // device spheres are read in place
  if (!is_device) {
    if ((rc = ensure(c, c->raybuf, 16 * (size_t)n))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->raybuf.p, spheres, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    d_s = reinterpret_cast<const float*>(c->raybuf.p);
  }
  // scan scratch for m values: [count m][offset m][block sums][total][bad]
  auto scratch = [&](uint32_t m) { return ensure(c, c->sg_scan, 2 * (size_t)m + (m + kScanBlock - 1) / kScanBlock + 2); };
  auto read_total = [&](uint32_t m, uint32_t* host, uint32_t words) -> int {
    const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
    return read_device_words(c, c->sg_scan.p + 2 * (size_t)m + nb, words, host);
  };
  auto scan = [&](uint32_t m) {
    const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
    uint32_t* cnt = c->sg_scan.p;
    return device_scan(c, cnt, cnt + m, m, cnt + 2 * (size_t)m, cnt + 2 * (size_t)m + nb);
  };
  // Spheres go in batches whose frontier stays within kMaxItems items; the
  // first batch is all of them, and a batch that outgrows it is halved.
  const uint32_t kMaxItems = c->sg_max_items;
  uint64_t total = 0;
  uint32_t s0 = 0, batch = n;
  while (s0 < n) {
    const uint32_t sn = std::min(batch, n - s0);
    const float* bs = d_s + 4 * (size_t)s0;
    if ((rc = scratch(sn))) return rc;
    dispatch_K<SgSeedCount>(c->K, c, body, sn, bs, c->sg_scan.p);
    HIPCHK(c, hipGetLastError());
    if ((rc = scan(sn))) return rc;
    uint32_t m = 0;
    if ((rc = read_total(sn, &m, 1))) return rc;
    bool too_big = m > kMaxItems;
    if (m && !too_big) {
      if ((rc = ensure(c, c->sg_items[0], m))) return rc;
      hipLaunchKernelGGL(k_sg_seed_write, dim3((sn + 255) / 256), dim3(256), 0, c->stream, c->dbodies.p, body, c->dtops.p, sn,
                         (const uint32_t*)c->sg_scan.p, (const uint32_t*)(c->sg_scan.p + sn), c->sg_items[0].p);
      HIPCHK(c, hipGetLastError());
    }
    int cur = 0;
    for (uint32_t step = 0; !too_big && m && step < steps; ++step) {
      if ((rc = scratch(m))) return rc;
      if ((rc = ensure(c, c->sg_kind, m))) return rc;
      dispatch_K<SgExpand2Count>(c->K, c, (const uint2*)c->sg_items[cur].p, m, bs, c->sg_scan.p, c->sg_kind.p);
      HIPCHK(c, hipGetLastError());
      if ((rc = scan(m))) return rc;
      uint32_t m2 = 0;
      if ((rc = read_total(m, &m2, 1))) return rc;
      if (m2 > kMaxItems) { too_big = true; break; }
      if (m2) {
        if ((rc = ensure(c, c->sg_items[cur ^ 1], m2))) return rc;
        dispatch_K<SgExpand2Write>(c->K, c, (const uint2*)c->sg_items[cur].p, m, (const uint8_t*)c->sg_kind.p,
                                   (const uint32_t*)(c->sg_scan.p + m), c->sg_items[cur ^ 1].p);
        HIPCHK(c, hipGetLastError());
      }
      cur ^= 1;
      m = m2;
    }
    if (too_big) {
      if (sn == 1) return fail(c, KDOP_ERR_CAPACITY, "kdop_single_traversal: sphere %u overlaps more than %u nodes of one level", s0, kMaxItems);
      batch = std::max(1u, sn / 2);
      continue;
    }
    if (m) {
      if ((rc = scratch(m))) return rc;
      const uint32_t nb = (m + kScanBlock - 1) / kScanBlock;
      uint32_t* bad = c->sg_scan.p + 2 * (size_t)m + nb + 1;
      HIPCHK(c, hipMemsetAsync(bad, 0, 4, c->stream));
      dispatch_K<SgLeaves>(c->K, c, body, (const uint2*)c->sg_items[cur].p, m, bs, s0, flip, 0, c->sg_scan.p, (const uint32_t*)nullptr,
                           (kdop_sphere_contact*)nullptr, bad);
      HIPCHK(c, hipGetLastError());
      if ((rc = scan(m))) return rc;
      uint32_t tb[2] = {0, 0};  // contacts of the batch, bad
      if ((rc = read_total(m, tb, 2))) return rc;
      if (tb[1]) return fail(c, KDOP_ERR_INTERNAL, "kdop_single_traversal: a node other than a leaf after %u steps", steps);
      if (tb[0]) {
        if (total + tb[0] >= (1ull << 32)) return fail(c, KDOP_ERR_CAPACITY, "kdop_single_traversal: more than 2^32 contacts");
        if ((rc = ensure(c, c->sphere_out, total + tb[0], true))) return rc;
        // the flagged leaves in order (the other frontier buffer holds the list), then their contacts
        if ((rc = ensure(c, c->sg_items[cur ^ 1], ((size_t)tb[0] + 1) / 2))) return rc;
        uint32_t* list = reinterpret_cast<uint32_t*>(c->sg_items[cur ^ 1].p);
        hipLaunchKernelGGL(k_gather_flagged, dim3((m + 255) / 256), dim3(256), 0, c->stream, (const uint32_t*)c->sg_scan.p,
                           (const uint32_t*)(c->sg_scan.p + m), m, list);
        dispatch_K<SgLeaves>(c->K, c, body, (const uint2*)c->sg_items[cur].p, tb[0], bs, s0, flip, 1, (uint32_t*)nullptr,
                             (const uint32_t*)list, c->sphere_out.p + total, (uint32_t*)nullptr);
        HIPCHK(c, hipGetLastError());
        total += tb[0];
      }
    }
    s0 += sn;
  }
  *count = total;
  if (!out) return KDOP_OK;
  if (cap < total) return fail(c, KDOP_ERR_CAPACITY, "kdop_single_traversal: need %llu contacts", (unsigned long long)total);
  if (total == 0) return KDOP_OK;
  HIPCHK(c, hipMemcpyAsync(out, c->sphere_out.p, (size_t)total * sizeof(kdop_sphere_contact),
                           is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KDOP_OK;
}

}  // extern "C"
