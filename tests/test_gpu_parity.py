"""Parity of the HIP pipeline (device build + refit + traversal + narrow phase)
against the CPU oracle, through the C ABI.

Bar: bit-exact for the tree volumes, world coordinates, leaf-pair sets and
test counts; contact points / normals / depths within 1e-5 relative (they
are expected bit-exact as well; the number of inexact values is asserted 0
for the default workloads).
"""
import numpy as np
import pytest

import kdop_amd as K
from kdop_amd import workloads as W
from helpers import (device_run, oracle_contacts_with_ordinals, oracle_run, sort_pairs)

pytestmark = pytest.mark.gpu

CONTACT_RTOL = 1e-5


def _check_pair_results(wl, dev, orc, exact_contacts=True):
    cnt = dev["counters"]
    tot = {k: 0 for k in ("tests", "descent_tests", "branch_overlaps", "leaf_pairs", "surface_pairs", "contacts",
                          "narrow_pairs")}
    p1_o, p2_o, con_o = [], [], []
    for pi, (c, p1, p2, con) in enumerate(orc):
        for k in tot:
            tot[k] += c[k]
        p1_o.append(np.column_stack([np.full(p1.shape[0], pi, np.uint32), p1]))
        p2_o.append(np.column_stack([np.full(p2.shape[0], pi, np.uint32), p2]))
        if con is not None and con.shape[0]:
            con_o.append((pi, con))
    # node-pair test count of the reference traversal (kdop_tandem_traversal.h)
    assert cnt["reference_tests"] == tot["tests"], (cnt, tot)
    assert cnt["descent_tests"] == tot["descent_tests"]
    assert cnt["branch_overlaps"] == tot["branch_overlaps"]
    assert cnt["leaf_pairs"] == tot["leaf_pairs"]
    assert cnt["surface_pairs"] == tot["surface_pairs"]
    # pair sets, bit exact after sorting
    p1_o = sort_pairs(np.concatenate(p1_o)) if p1_o else np.zeros((0, 3), np.uint32)
    p2_o = sort_pairs(np.concatenate(p2_o)) if p2_o else np.zeros((0, 3), np.uint32)
    if "p1" in dev:
        np.testing.assert_array_equal(sort_pairs(dev["p1"]), p1_o)
    np.testing.assert_array_equal(sort_pairs(dev["p2"]), p2_o)
    if "contacts" not in dev:
        return
    assert cnt["contacts"] == tot["contacts"]
    assert cnt["narrow_pairs"] == tot["narrow_pairs"]
    dc = dev["contacts"]
    rows = []
    for pi, con in con_o:
        ords = oracle_contacts_with_ordinals(con)
        for c, o in zip(con, ords):
            rows.append((pi, int(c["ta"]), int(c["tb"]), int(o), *c["p"], *c["n"], float(c["d"])))
    if not rows:
        assert dc.shape[0] == 0
        return
    ref = np.array(rows, dtype=np.float64)
    ref = ref[np.lexsort((ref[:, 3], ref[:, 2], ref[:, 1], ref[:, 0]))]
    key = np.lexsort((dc["ordinal"], dc["tet_b"], dc["tet_a"], dc["pair"]))
    dc = dc[key]
    np.testing.assert_array_equal(dc["pair"], ref[:, 0].astype(np.uint32))
    np.testing.assert_array_equal(dc["tet_a"], ref[:, 1].astype(np.uint32))
    np.testing.assert_array_equal(dc["tet_b"], ref[:, 2].astype(np.uint32))
    np.testing.assert_array_equal(dc["ordinal"], ref[:, 3].astype(np.uint32))
    got = np.concatenate([dc["p"], dc["n"], dc["depth"][:, None]], axis=1).astype(np.float64)
    want = ref[:, 4:]
    scale = np.maximum(np.abs(want), 1.0)
    assert np.all(np.abs(got - want) <= CONTACT_RTOL * scale)
    if exact_contacts:
        assert np.count_nonzero(got.astype(np.float32) != want.astype(np.float32)) == 0


MODES = [K.KDOP_TRAVERSAL_PERSISTENT, K.KDOP_TRAVERSAL_LEVELS]


@pytest.mark.parametrize("mode", MODES, ids=["persistent", "levels"])
def test_single_pair_box10_k8(mode):
    """Two copies of the 10^3-cube box (5000 tets), K=8: full parity."""
    wl = W.single_pair(n=10, K=8)
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl, traversal=mode)
    _check_pair_results(wl, dev, orc)
    assert dev["counters"]["leaf_pairs"] > 100000


@pytest.mark.parametrize("mode", MODES, ids=["persistent", "levels"])
@pytest.mark.parametrize("k", [6, 8, 12, 14, 18, 20, 26])
def test_jittered_pair_all_k(k, mode):
    """Jittered 12^3-cube boxes (8640 tets), random pose, every supported K."""
    wl = W.large_pair(n=12, K=k, seed=3)
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl, traversal=mode)
    _check_pair_results(wl, dev, orc)


def test_refit_bits_and_world_coordinates():
    """Device world coordinates and every tree volume equal the oracle's bits."""
    from helpers import world_coords
    wl = W.large_pair(n=8, K=26, seed=5)
    orc, trees, worlds = oracle_run(wl, narrow=False)
    dev = device_run(wl, narrow=False)
    ctx = dev["ctx"]
    for i, bid in enumerate(dev["ids"]):
        X, Y, Z = ctx.world(bid)
        np.testing.assert_array_equal(X, worlds[i][0])
        np.testing.assert_array_equal(Y, worlds[i][1])
        np.testing.assert_array_equal(Z, worlds[i][2])
        t = ctx.export_tree(bid)
        np.testing.assert_array_equal(t["root"], trees[i].root())
        ex = trees[i].export()
        ref = {}
        for (lv, ch, sl, leaf, st, en), (f, l), v in zip(ex["info"], ex["range"], ex["vol"]):
            ref[(int(f), int(l))] = v
        for (f, l), v in zip(t["range"], t["vol"]):
            np.testing.assert_array_equal(v, ref[(int(f), int(l))])


@pytest.mark.parametrize("mode", MODES, ids=["persistent", "levels"])
def test_batched_pairs_one_launch(mode):
    """Several body pairs of different sizes (some separated) in one call."""
    wl = W.batched_pairs(npairs=12, K=14, seed=4, tmin=200, tmax=6000)
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl, traversal=mode)
    _check_pair_results(wl, dev, orc)


@pytest.mark.parametrize("ls,wpc", [(256, 1), (512, 4), (2048, 16)])
def test_persistent_queue_shapes(ls, wpc):
    """Persistent kernel with one stack per wave, tiny and large LDS stacks /
    few and many waves: stack overflow into the next round must not change
    results."""
    wl = W.large_pair(n=10, K=18, seed=13)
    orc, _, _ = oracle_run(wl)
    ctx = K.Context(K=wl.K, traversal=K.KDOP_TRAVERSAL_PERSISTENT, queue_depth=ls, waves_per_cu=wpc)
    ctx.set_option("block_waves", 1)
    dev = device_run(wl, ctx=ctx)
    _check_pair_results(wl, dev, orc)


@pytest.mark.parametrize("bw,wpc,keep", [(2, 8, 512), (4, 8, 512), (8, 8, 512), (2, 2, 16), (4, 16, 64), (8, 4, 1)],
                         ids=["w2", "w4", "w8", "w2_few_waves_small_keep", "w4_many_waves", "w8_hand_off_all"])
def test_persistent_shared_stack_blocks(bw, wpc, keep):
    """Persistent rounds with bw waves per workgroup sharing one LDS stack:
    the oracle's counts, pair sets and contacts, also when a small keep hands
    most of each block's stack to the next round and with few or many waves."""
    for wl in (W.large_pair(n=10, K=18, seed=13), W.batched_pairs(npairs=10, K=26, seed=8, tmin=300, tmax=5000)):
        orc, _, _ = oracle_run(wl)
        ctx = K.Context(K=wl.K, traversal=K.KDOP_TRAVERSAL_PERSISTENT, waves_per_cu=wpc, keep=keep)
        ctx.set_option("block_waves", bw)
        assert ctx.get_option("block_waves") == bw
        dev = device_run(wl, ctx=ctx)
        print("bw", bw, "wpc", wpc, "keep", keep, "rounds", ctx.traversal_stats()["rounds"], flush=True)
        _check_pair_results(wl, dev, orc)


@pytest.mark.parametrize("bw,k", [(4, 26), (2, 18), (4, 14), (4, 8), (2, 6)])
def test_persistent_node_slabs(bw, k):
    """Persistent rounds that stage each wave's node records in an LDS node
    slab (KDOP_OPT_NODE_SLAB) before testing: the oracle's counts, pair sets
    and contacts, for every record width (K = 6, 8-14, 18-26) and both slab
    workgroup shapes, on a single pair and on a batch of uneven pairs."""
    for wl in (W.large_pair(n=10, K=k, seed=13), W.batched_pairs(npairs=10, K=k, seed=8, tmin=300, tmax=5000)):
        orc, _, _ = oracle_run(wl)
        ctx = K.Context(K=wl.K, traversal=K.KDOP_TRAVERSAL_PERSISTENT)
        ctx.set_option("block_waves", bw)
        ctx.set_option("node_slab", 1)
        assert ctx.get_option("node_slab") == 1
        dev = device_run(wl, ctx=ctx)
        _check_pair_results(wl, dev, orc)


def test_duplicate_body_pairs_fall_back_to_levels():
    """The same body pair twice in one call: both copies are reported."""
    wl = W.single_pair(n=5, K=8)
    wl.pairs = np.array([[0, 1], [0, 1]], np.uint32)
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl)
    _check_pair_results(wl, dev, orc)


def test_separated_bodies():
    """Bodies far apart: one root test per pair, nothing else."""
    wl = W.single_pair(n=4, K=8)
    wl.poses[1] = W.Pose(np.array([5.0, 0.0, 0.0], np.float32), wl.poses[1].q)
    dev = device_run(wl)
    c = dev["counters"]
    assert c["tests"] == 1 and c["reference_tests"] == 1
    assert c["leaf_pairs"] == 0 and dev["p2"].shape[0] == 0 and dev["contacts"].shape[0] == 0


def test_refit_series_small():
    """Three steps of rigid motion with refit (not rebuild): pair sets per step."""
    wl = W.large_pair(n=8, K=18, seed=9)
    ctx = K.Context(K=18)
    ids = [ctx.make_tree(m) for m in wl.meshes]
    P, Q = W.motion(2, 3, seed=31, dt=0.05, drift=0.01)
    from oracle import oracle as O
    from helpers import world_coords
    trees = [O.Tree(18, 8000, m.tets, m.X, m.Y, m.Z) for m in wl.meshes]
    for s in range(3):
        poses = [W.Pose(P[s, i] + wl.poses[i].p, Q[s, i]) for i in range(2)]
        ctx.refit(ids, np.stack([p.p for p in poses]), np.stack([p.q for p in poses]))
        cnt = ctx.tandem_traversal([[ids[0], ids[1]]], narrow=True, keep_all_pairs=True)
        ws = [world_coords(m, p) for m, p in zip(wl.meshes, poses)]
        for t, w in zip(trees, ws):
            t.refit(*w)
        c, p1, p2, con, _ = O.traverse(trees[0], (wl.meshes[0].tets, *ws[0], wl.meshes[0].surface),
                                       trees[1], (wl.meshes[1].tets, *ws[1], wl.meshes[1].surface))
        assert cnt["reference_tests"] == c["tests"]
        np.testing.assert_array_equal(sort_pairs(ctx.pairs(0)[:, 1:]), sort_pairs(p1))
        assert cnt["contacts"] == c["contacts"]


def test_autotuned_batched_parity():
    """The autotuner's final run (its chosen options) reports the oracle's
    results, and the options it leaves on the context are valid ones."""
    wl = W.batched_pairs(npairs=10, K=8, seed=14, tmin=300, tmax=4000)
    orc, _, _ = oracle_run(wl)
    ctx = K.Context(K=8)
    ids = [ctx.make_tree(m) for m in wl.meshes]
    ctx.refit(ids, np.stack([p.p for p in wl.poses]), np.stack([p.q for p in wl.poses]))
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    chosen, best_ms = ctx.autotune(pairs, narrow=True, budget_ms=1500)
    assert best_ms > 0 and chosen["queue_depth"] in (128, 256, 512, 1024, 2048) and 1 <= chosen["waves_per_cu"] <= 32
    dev = device_run(wl, ctx=ctx, ids=ids)
    _check_pair_results(wl, dev, orc)


def test_async_refit_graph_replay_and_state_reuse():
    """Refits that return once queued and traversals replayed from the
    context's HIP graph give, call after call, the results of synchronous
    refits and direct launches (the per-call state is cleared by the previous
    call's final fold, not by a memset); every replay is checked."""
    wl = W.large_pair(n=10, K=18, seed=23)
    P, Q = W.motion(2, 6, seed=5, dt=0.05, drift=0.01)
    ctxs = [K.Context(K=18), K.Context(K=18)]
    ctxs[0].set_option("graphs", 0)
    ctxs[1].set_option("graphs", 1)
    ctxs[1].set_option("async_refit", 1)
    ids = [[c.make_tree(m) for m in wl.meshes] for c in ctxs]
    for s in range(6):
        pos = np.stack([P[s, i] + wl.poses[i].p for i in range(2)])
        quat = Q[s]
        res = []
        for c, b in zip(ctxs, ids):
            c.refit(b, pos, quat)
            cnt = c.tandem_traversal([[b[0], b[1]]], narrow=True, keep_all_pairs=True)
            res.append((cnt, sort_pairs(c.pairs(0)), c.contacts()))
        assert res[0][0] == res[1][0]
        np.testing.assert_array_equal(res[0][1], res[1][1])
        key = lambda con: np.lexsort((con["ordinal"], con["tet_b"], con["tet_a"]))  # noqa: E731
        np.testing.assert_array_equal(res[0][2][key(res[0][2])], res[1][2][key(res[1][2])])


@pytest.mark.parametrize("path", [1, 2], ids=["wave_per_pair", "gate_then_contacts"])
def test_both_narrow_paths(path):
    """The narrow phase's two paths (chosen by the surface-pair count) give
    the oracle's contacts: forced one-pass and two-pass on the same input."""
    wl = W.large_pair(n=12, K=14, seed=19)
    orc, _, _ = oracle_run(wl)
    ctx = K.Context(K=14)
    ctx.set_option("narrow_path", path)
    dev = device_run(wl, ctx=ctx)
    _check_pair_results(wl, dev, orc)


def test_levels_first_call_regrows_segmented_frontier():
    """A fresh context starts with a 65,536-item frontier; a level-synchronous
    call whose levels hold millions of items overflows its segments, grows the
    buffers and runs again (counters['overflow'] > 0) with full parity; the
    next call on the grown buffers does not overflow and gives the same sets."""
    # a body against an identical copy in the same pose: every chunk overlaps
    # its neighbours, so the levels far exceed the first estimate (4 per tet)
    mesh = K.box_mesh(14, size=1.0, jitter=0.1, seed=6)
    wl = W.PairWorkload("self14_k8", 8, [mesh, mesh], np.array([[0, 1]], np.uint32), [W.IDENTITY, W.IDENTITY])
    orc, _, _ = oracle_run(wl)
    dev = device_run(wl, traversal=K.KDOP_TRAVERSAL_LEVELS)
    print("overflow", dev["counters"]["overflow"], "tests", dev["counters"]["tests"])
    assert dev["counters"]["overflow"] > 0
    _check_pair_results(wl, dev, orc)
    again = device_run(wl, ctx=dev["ctx"], ids=dev["ids"])
    assert again["counters"]["overflow"] == 0
    np.testing.assert_array_equal(sort_pairs(again["p1"]), sort_pairs(dev["p1"]))
    np.testing.assert_array_equal(sort_pairs(again["p2"]), sort_pairs(dev["p2"]))


def test_queued_rounds_follow_recent_calls():
    """The persistent strategy queues as many rounds as recent calls needed.
    One context alternates between a pair that needs one round and calls
    that spill over several (a stack keep of 1): every call, whether its
    rounds were queued or drained one by one, gives the level-synchronous
    strategy's counts, surface pairs and contacts, including the calls
    right after the number of rounds changed and graph replays.  The
    spilling calls take a round budget of one iteration, so every block
    hands its stack on after its first iteration without frontier and those
    calls run many rounds whatever the timing; the others take no budget and
    keep their stacks (one round, or two when a stack overflows)."""
    wl = W.batched_pairs(npairs=16, K=14, seed=21, tmin=2000, tmax=40000)
    key = lambda a: sorted(a[i:i + 1].tobytes() for i in range(a.shape[0]))
    ref = device_run(wl, keep_all=False, traversal=K.KDOP_TRAVERSAL_LEVELS)
    want = (ref["counters"], key(ref["p2"]), key(ref["ctx"].contacts()))
    ctx = K.Context(K=wl.K, traversal=K.KDOP_TRAVERSAL_PERSISTENT)
    ctx.set_option("graphs", 1)
    ids = [ctx.make_tree(m) for m in wl.meshes]
    ctx.refit(ids, np.stack([p.p for p in wl.poses]), np.stack([p.q for p in wl.poses]))
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    seen = {}
    for keep, iters in [(512, 0)] * 12 + [(1, 1)] * 3 + [(512, 0)] * 12:
        ctx.set_option("stack_keep", keep)
        ctx.set_option("round_iters", iters)
        c = ctx.tandem_traversal(pairs, narrow=True)
        seen.setdefault(keep, set()).add(ctx.traversal_stats()["rounds"])
        for f in ("tests", "leaf_pairs", "surface_pairs", "contacts", "narrow_pairs", "overflow"):
            assert c[f] == want[0][f], (keep, f)
        assert key(ctx.pairs(1)) == want[1]
        assert key(ctx.contacts()) == want[2]
    print("rounds seen", seen, flush=True)
    # calls after the spilling ones queue as many rounds as those needed
    # until they see fewer are enough, so only the smallest counts compare
    assert min(seen[512]) <= 2 < min(seen[1])


def test_ramp_stops_where_the_last_call_stopped():
    """A call from the same seeds queues no breadth-first launches past the
    level where the previous call's ramp found a target-wide frontier (they
    would only find the stop flag); every call still gives the oracle's
    results, and so do calls after the target changes (the history no longer
    applies) and after the pair list changes (new seeds)."""
    wl = W.batched_pairs(npairs=12, K=14, seed=4, tmin=200, tmax=6000)
    orc, _, _ = oracle_run(wl)
    ctx = K.Context(K=14, traversal=K.KDOP_TRAVERSAL_PERSISTENT)
    ctx.set_option("bfs_max", 6)
    ids = None
    stops = []
    for target in (256, 256, 256, 4096, 4096):
        ctx.set_option("ramp_target", target)
        dev = device_run(wl, ctx=ctx, ids=ids)
        ids = dev["ids"]
        stops.append(ctx.traversal_stats()["ramp_levels"])
        _check_pair_results(wl, dev, orc)
    print("ramp stop levels", stops, flush=True)
    # the same seeds and target stop at the same level, whether the levels
    # after it were queued (call 1) or not (calls 2, 3)
    assert stops[0] == stops[1] == stops[2] < 6 and stops[3] == stops[4] >= stops[0]
    # a shorter pair list (new seeds) from the same context
    sub = W.PairWorkload(wl.name, wl.K, wl.meshes, wl.pairs[:5], wl.poses)
    orc5 = orc[:5]
    dev = device_run(sub, ctx=ctx, ids=ids)
    _check_pair_results(sub, dev, orc5)


def test_seed_budget_follows_the_pair_count():
    """An unset seed budget is 262,144 chunk-root pairs for a call of 256
    pairs or more and 65,536 below (the option reports the last call's); a
    budget the caller sets holds for every call until it is set to 0.  Every
    call gives the oracle's results."""
    wl = W.batched_pairs(npairs=256, K=8, seed=11, tmin=20, tmax=400)
    orc, _, _ = oracle_run(wl)
    one = W.PairWorkload(wl.name, wl.K, wl.meshes, wl.pairs[:1], wl.poses)
    ctx = K.Context(K=8, traversal=K.KDOP_TRAVERSAL_PERSISTENT)
    dev = device_run(wl, ctx=ctx)
    ids = dev["ids"]
    _check_pair_results(wl, dev, orc)
    assert ctx.get_option("seed_budget") == 262144
    dev = device_run(one, ctx=ctx, ids=ids)
    _check_pair_results(one, dev, orc[:1])
    assert ctx.get_option("seed_budget") == 65536
    ctx.set_option("seed_budget", 4096)
    dev = device_run(wl, ctx=ctx, ids=ids)
    _check_pair_results(wl, dev, orc)
    assert ctx.get_option("seed_budget") == 4096
    ctx.set_option("seed_budget", 0)
    dev = device_run(wl, ctx=ctx, ids=ids)
    _check_pair_results(wl, dev, orc)
    assert ctx.get_option("seed_budget") == 262144
