"""Parity at the benchmark configurations' real sizes, and edge cases.

Every comparison is against the CPU oracle on the same inputs, to the same
bar as tests/test_gpu_parity.py: bit-exact pair sets (after sorting), node-pair
test counts and contact values.
"""
import numpy as np
import pytest

import kdop_amd as K
from kdop_amd import workloads as W
from helpers import device_run, oracle_run, sort_pairs, world_coords
from oracle import oracle as O
from test_gpu_parity import _check_pair_results

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 8


@pytest.mark.parametrize("k", [26, 6, 14, 18])
def test_c2_full_size(k):
    """Two jittered 40^3 boxes (320,000 tets each), the headline pose: full
    parity (P1, P2, counts incl. the reference's 1.6e9-test count at K=26,
    contacts) for the headline K and the comparison K values."""
    wl = W.large_pair(n=40, K=k, seed=7)
    orc, _, _ = oracle_run(wl, threads=ORACLE_THREADS)
    dev = device_run(wl)
    _check_pair_results(wl, dev, orc)
    if k == 26:
        assert dev["counters"]["reference_tests"] == 1603611489


def test_c3_batched_subset_full_sizes():
    """128 pairs drawn like the batched benchmark (1k-100k tets, ~40 %
    separated, 30-60 % AABB overlap), one call, both traversal strategies."""
    wl = W.batched_pairs(npairs=128, K=8, seed=12)
    orc, _, _ = oracle_run(wl, threads=ORACLE_THREADS)
    for mode in (K.KDOP_TRAVERSAL_PERSISTENT, K.KDOP_TRAVERSAL_LEVELS):
        dev = device_run(wl, keep_all=False, traversal=mode)
        _check_pair_results(wl, dev, orc)


def test_c3_small_stack_keep_completes():
    """The full 512-pair batched draw under the persistent strategy with a
    stack keep of 1 and 32 (most of every stack handed to the next round),
    one stack per wave and two waves per stack: the call completes (the
    rounds drained after the queued ones hand nothing on) and gives the
    level-synchronous strategy's counts, surface pairs and contacts."""
    wl = W.batched_pairs(npairs=512, K=8, seed=11)
    ref = device_run(wl, keep_all=False, traversal=K.KDOP_TRAVERSAL_LEVELS)
    rc = ref["counters"]
    key = lambda a: sorted(a[i:i + 1].tobytes() for i in range(a.shape[0]))
    ref_p2, ref_c = key(ref["p2"]), key(ref["ctx"].contacts())
    for bw, keep in ((1, 1), (2, 1), (2, 32)):
        ctx = K.Context(K=8, traversal=K.KDOP_TRAVERSAL_PERSISTENT, keep=keep)
        ctx.set_option("block_waves", bw)
        dev = device_run(wl, keep_all=False, ctx=ctx)
        c = dev["counters"]
        print("bw", bw, "keep", keep, "rounds", ctx.traversal_stats()["rounds"], "tests", c["tests"], flush=True)
        for f in ("tests", "leaf_pairs", "surface_pairs", "contacts", "narrow_pairs", "overflow"):
            assert c[f] == rc[f], f
        assert key(dev["p2"]) == ref_p2
        assert key(ctx.contacts()) == ref_c
        ctx.close()


def test_c3_wide_iterations_claim_frontier():
    """The full 512-pair batched draw under the persistent strategy with
    workgroups that expand 128 items per iteration (eight waves per stack,
    or four waves with two items per quad): the frontier is wide enough that
    a workgroup would claim more frontier items than one wave has lanes, and
    the claims are capped at one item per lane of the wave that loads them.
    Counts, surface pairs and contacts equal the level-synchronous
    strategy's."""
    wl = W.batched_pairs(npairs=512, K=8, seed=11)
    ref = device_run(wl, keep_all=False, traversal=K.KDOP_TRAVERSAL_LEVELS)
    rc = ref["counters"]
    key = lambda a: sorted(a[i:i + 1].tobytes() for i in range(a.shape[0]))
    ref_p2, ref_c = key(ref["p2"]), key(ref["ctx"].contacts())
    for opts in ({"block_waves": 8}, {"block_waves": 4, "quad_items": 2}, {"block_waves": 2}):
        ctx = K.Context(K=8, traversal=K.KDOP_TRAVERSAL_PERSISTENT)
        for name, v in opts.items():
            ctx.set_option(name, v)
        dev = device_run(wl, keep_all=False, ctx=ctx)
        c = dev["counters"]
        st = ctx.traversal_stats()
        print(opts, "rounds", st["rounds"], "grabs", st["grabs"], "tests", c["tests"], flush=True)
        for f in ("tests", "leaf_pairs", "surface_pairs", "contacts", "narrow_pairs", "overflow"):
            assert c[f] == rc[f], f
        assert key(dev["p2"]) == ref_p2
        assert key(ctx.contacts()) == ref_c
        ctx.close()


def test_self_contact_completes():
    """A jittered 20^3 box (40,000 tets, K=26) against an identical copy at
    identity, the densest traversal these meshes make: under AUTO, under the
    persistent strategy with a stack keep of 1 (the queued rounds hand the
    bottom of every deep stack on), and with one persistent round only and a
    keep of 1 (the hand-offs spill into a round that is not run, so the
    spill is finished by the level-synchronous launches), with either item
    mapping, every call completes with the oracle's counts, pair sets and
    contacts."""
    wl = W.self_contact(n=20, K=26, seed=3)
    orc, _, _ = oracle_run(wl, threads=ORACLE_THREADS)
    cases = [
        (dict(traversal=K.KDOP_TRAVERSAL_AUTO), {}),
        (dict(traversal=K.KDOP_TRAVERSAL_PERSISTENT, keep=1), {}),
        (dict(traversal=K.KDOP_TRAVERSAL_PERSISTENT, keep=1), {"item_lanes": 1}),
        (dict(traversal=K.KDOP_TRAVERSAL_PERSISTENT, keep=1, waves_per_cu=4), {"max_rounds": 1, "block_waves": 2}),
        (dict(traversal=K.KDOP_TRAVERSAL_PERSISTENT, keep=1, queue_depth=256, waves_per_cu=2),
         {"max_rounds": 1, "item_lanes": 1, "block_waves": 1}),
    ]
    for kw, opts in cases:
        ctx = K.Context(K=26, **kw)
        for name, v in opts.items():
            ctx.set_option(name, v)
        dev = device_run(wl, ctx=ctx)
        st = ctx.traversal_stats()
        print(kw, opts, "rounds", st["rounds"], "handoffs", st["handoffs"], "fallback levels", st["fallback_levels"],
              "tests", dev["counters"]["tests"], flush=True)
        _check_pair_results(wl, dev, orc)
        if kw.get("keep") == 1:
            assert st["handoffs"] > 0
        if "max_rounds" in opts:
            assert st["fallback_levels"] > 0 and st["rounds"] == opts["max_rounds"]
        ctx.close()


SIZE_MAX = 2 ** 64 - 1


@pytest.mark.parametrize("k,mem_a,mem_b", [(8, 16000, 16000), (8, 32000, 32000), (8, 32000, 8000), (26, 131072, 16000),
                                           (14, SIZE_MAX, SIZE_MAX), (18, SIZE_MAX, 8000)],
                         ids=["k8_L64", "k8_L128", "k8_L128_vs_L32", "k26_L256_vs_L16", "k14_one_chunk", "k18_one_chunk_vs_L16"])
def test_per_body_mem_bytes(k, mem_a, mem_b):
    """kdop_add_body_ex with each body's own mem_bytes (make_tree's first
    argument, kdop_make_tree.h:190-229): chunks of 64 to 256 leaves, one chunk
    of every tet (SIZE_MAX, the gProximity setting), and two bodies of one
    context built with different values.  Tree shape, every tree volume, the
    counts, pair sets and contacts equal the oracle's built with the same
    mem_bytes, over a refit to a new pose and in both traversal strategies."""
    wl = W.large_pair(n=10, K=k, seed=13)
    orc, trees, worlds = oracle_run(wl, chunk_bytes=[mem_a, mem_b], threads=ORACLE_THREADS)
    for mode in (K.KDOP_TRAVERSAL_PERSISTENT, K.KDOP_TRAVERSAL_LEVELS):
        ctx = K.Context(K=k, traversal=mode)
        ids = [ctx.make_tree(m, mem_bytes=mb) for m, mb in zip(wl.meshes, (mem_a, mem_b))]
        for bid, mb, t in zip(ids, (mem_a, mem_b), trees):
            sh, ref = ctx.shape(bid), K.tree_shape(k, mb, wl.meshes[0].nt)
            assert (sh["L"], sh["C"], sh["H"]) == (ref["L"], ref["C"], ref["H"])
            assert len(t.levels()) == ref["H"] and t.levels()[-1] == ref["C"]
        dev = device_run(wl, ctx=ctx, ids=ids)
        print(k, mem_a, mem_b, "mode", mode, "L", ctx.shape(ids[0])["L"], ctx.shape(ids[1])["L"], "tests",
              dev["counters"]["tests"], flush=True)
        _check_pair_results(wl, dev, orc)
        for i, bid in enumerate(ids):
            ex, te = trees[i].export(), ctx.export_tree(bid)
            np.testing.assert_array_equal(te["root"], trees[i].root())
            ref = {(int(f), int(l)): v for (f, l), v in zip(ex["range"], ex["vol"])}
            for (f, l), v in zip(te["range"], te["vol"]):
                np.testing.assert_array_equal(v, ref[(int(f), int(l))])
        ctx.close()


def test_c4_refit_series():
    """64 overlapping pairs (20k-80k tets, K=18), 200 steps of rigid motion,
    refit (never rebuilt) and re-traversed every step on the device; the
    pair lists, counts and contacts of steps 0, 7, 40 and 199 equal the
    oracle's, and every step reports a consistent count set."""
    wl, P, Q = W.refit_series(npairs=64, K=18, steps=200)
    ctx = K.Context(K=18)
    ids = np.array([ctx.make_tree(m) for m in wl.meshes], np.uint32)
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    check = {0, 7, 40, 199}
    for s in range(P.shape[0]):
        ctx.refit(ids, P[s], Q[s])
        cnt = ctx.tandem_traversal(pairs, narrow=True, keep_all_pairs=s in check)
        assert cnt["surface_pairs"] <= cnt["leaf_pairs"] and cnt["narrow_pairs"] <= cnt["surface_pairs"]
        if s not in check:
            continue
        wl.poses = [W.Pose(P[s, i], Q[s, i]) for i in range(len(wl.meshes))]
        orc, _, _ = oracle_run(wl, threads=ORACLE_THREADS)
        dev = dict(counters=cnt, p1=ctx.pairs(0), p2=ctx.pairs(1), contacts=ctx.contacts())
        _check_pair_results(wl, dev, orc)


def _tet_mesh(points, tets):
    pts = np.asarray(points, np.float32)
    T = np.asarray(tets, np.uint32).reshape(-1, 4)
    return K.Mesh(pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), T, K.compute_surface_map(pts.shape[0], T))


def test_empty_call_and_single_tet_bodies():
    """No pairs at all; then two one-tet bodies (L = 1, one chunk, one level)
    that overlap, that touch only by translation, and one against itself."""
    ctx = K.Context(K=8)
    assert ctx.tandem_traversal(np.zeros((0, 2), np.uint32))["tests"] == 0
    assert ctx.pairs(1).shape[0] == 0 and ctx.contacts().shape[0] == 0
    tet = _tet_mesh([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], [[0, 1, 2, 3]])
    for dp in ([0.1, 0.1, 0.1], [0.3, 0.0, 0.0], [5.0, 0.0, 0.0]):
        wl = W.PairWorkload("tets", 8, [tet, tet], np.array([[0, 1]], np.uint32),
                            [W.IDENTITY, W.Pose(np.array(dp, np.float32), np.array([0, 0, 0, 1], np.float32))])
        orc, _, _ = oracle_run(wl)
        _check_pair_results(wl, device_run(wl), orc)
    wl = W.PairWorkload("self", 8, [tet], np.array([[0, 0]], np.uint32), [W.IDENTITY])
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl)
    _check_pair_results(wl, dev, orc)
    assert dev["counters"]["leaf_pairs"] == 1


@pytest.mark.parametrize("k", [8, 26])
def test_chunk_boundary_tet_counts(k):
    """Bodies whose tet counts sit on chunk boundaries (L - 1, L, L + 1,
    2L - 1, 2L + 1 for the body's own L), cut from a jittered box in Morton
    order: partial and full last chunks, one and two chunk levels."""
    base = K.box_mesh(6, size=1.0, jitter=0.1, seed=3)
    meshes = []
    for nt in (7, 8, 9, 15, 17, 31, 32, 33, 63, 65, 500, 513):
        T = base.tets[:nt]
        meshes.append(K.Mesh(base.X, base.Y, base.Z, T, K.compute_surface_map(base.nv, T)))
    q = K.rng_quaternions(5, 0, len(meshes))
    poses = [W.Pose(np.array([0.02 * i, 0.01, 0.0], np.float32), q[i]) for i in range(len(meshes))]
    pairs = np.array([[i, (i + 1) % len(meshes)] for i in range(len(meshes))], np.uint32)
    wl = W.PairWorkload("boundary", k, meshes, pairs, poses)
    orc, trees, worlds = oracle_run(wl)
    dev = device_run(wl)
    _check_pair_results(wl, dev, orc)
    ctx = dev["ctx"]
    for i, bid in enumerate(dev["ids"]):   # every tree volume, bit for bit
        t = ctx.export_tree(bid)
        np.testing.assert_array_equal(t["root"], trees[i].root())
        ref = {(int(f), int(l)): v for (f, l), v in zip(trees[i].export()["range"], trees[i].export()["vol"])}
        for (f, l), v in zip(t["range"], t["vol"]):
            np.testing.assert_array_equal(v, ref[(int(f), int(l))])


@pytest.mark.parametrize("k,chunk_bytes", [(26, 2000), (26, 4000), (26, 8000), (14, 8000), (8, 8000), (6, 2000), (8, 2000),
                                           (12, 4000), (20, 4000), (20, 8000)],
                         ids=["L2", "L4", "L8", "L16", "L32", "k6_L8", "k8_L8", "k12_L8", "k20_L8", "k20_L16"])
def test_refit_tree_bits_around_super_chunks(k, chunk_bytes):
    """Tree volumes bit for bit for every chunk width L, on tet counts around
    L and L*L: level H-2 is refitted together with the leaves where L*L <= 64
    (whole super chunks in a wave) or L = 16 (one per block), level H-3 too
    where every body has L = 8 (512-thread blocks), and by the super-chunk
    pass where a chunk or super chunk is not full.  Every record width (32,
    64 and 128 bytes, meta words starting a 16-byte word or ending one)."""
    L = K.tree_shape(k, chunk_bytes, 4096)["L"]
    base = K.box_mesh(9, size=1.0, jitter=0.1, seed=11)
    counts = sorted({min(c, base.tets.shape[0]) for c in
                     (L * L - 1, L * L, L * L + 1, 2 * L * L + L - 1, 3 * L * L + 1, 4 * L * L + L + 1, 64 * L + 3, 1000, 3645)})
    meshes = []
    for nt in counts:
        T = base.tets[:nt]
        meshes.append(K.Mesh(base.X, base.Y, base.Z, T, K.compute_surface_map(base.nv, T)))
    q = K.rng_quaternions(7, 0, len(meshes))
    poses = [W.Pose(np.array([0.01 * i, 0.02, 0.0], np.float32), q[i]) for i in range(len(meshes))]
    pairs = np.array([[i, (i + 1) % len(meshes)] for i in range(len(meshes))], np.uint32)
    wl = W.PairWorkload("superchunks", k, meshes, pairs, poses)
    orc, trees, _ = oracle_run(wl, chunk_bytes=chunk_bytes)
    dev = device_run(wl, chunk_bytes=chunk_bytes)
    _check_pair_results(wl, dev, orc)
    ctx = dev["ctx"]
    for i, bid in enumerate(dev["ids"]):
        t = ctx.export_tree(bid)
        np.testing.assert_array_equal(t["root"], trees[i].root())
        ex = trees[i].export()
        ref = {(int(f), int(l)): v for (f, l), v in zip(ex["range"], ex["vol"])}
        assert len(ref) == len(t["range"])
        for (f, l), v in zip(t["range"], t["vol"]):
            np.testing.assert_array_equal(v, ref[(int(f), int(l))])


def test_invalid_inputs_fail_loudly():
    """Unknown body ids and out-of-range options are errors, not crashes."""
    ctx = K.Context(K=8)
    m = K.box_mesh(2)
    a = ctx.make_tree(m)
    with pytest.raises(K.KdopError):
        ctx.tandem_traversal([[a, a + 5]])
    with pytest.raises(K.KdopError):
        ctx.set_option("queue_depth", 300)
    with pytest.raises(K.KdopError):
        ctx.narrow_adjoint([a])   # no narrow traversal yet
