#!/usr/bin/env python3
"""Benchmark of the MI355X kDOP collision pipeline.

One step = one collision query of the headline workload, with the meshes and
trees already resident in HBM: rigid transform + refit of both bodies, tandem
traversal, tet-tet narrow phase (MOST_OPPOSING_SURFACES) on the surface pairs.

Headline workload: one pair of jittered 40^3-cube boxes (320,000 tets each),
K = 26, seeded pose with 30-60 % AABB overlap.  The JSON line reports
node-pair tests/s executed by the device over the whole step, the roofline of
the traversal kernel (algorithmic bytes per node-pair test = 2 (4K + 8)), the
batched leg (512 body pairs of 1k-100k tets in one call vs 512 calls), and the
CPU oracle timed on a bounded single-threaded sample of the same query.

    python bench.py [--steps 20] [--warmup 3] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bvh-tvcg18_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def rec_stride(K):
    """Floats per node record: the power of two >= K + 2 (8, 16 or 32)."""
    return 8 if K + 2 <= 8 else (16 if K + 2 <= 16 else 32)
METRIC = "BVTT node-pair tests/s + batched pair-queries/s; % HBM roofline; vs CPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_pair(ctx, wl):
    import numpy as np
    ids = [ctx.make_tree(m) for m in wl.meshes]
    pose_p = np.stack([p.p for p in wl.poses]).astype(np.float32)
    pose_q = np.stack([p.q for p in wl.poses]).astype(np.float32)
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    return ids, pose_p, pose_q, pairs


def device_poses(pose_p, pose_q):
    """Pose arrays in device memory (torch), so the timed steps read inputs from HBM."""
    import torch
    p = torch.from_numpy(pose_p).to("cuda")
    q = torch.from_numpy(pose_q).to("cuda")
    torch.cuda.synchronize()
    return p, q


def headline_leg(args):
    import numpy as np
    import torch
    import kdop_amd as KA
    from kdop_amd import workloads as W
    wl = W.large_pair(n=args.n, K=args.k, seed=7)
    ctx = KA.Context(K=args.k, traversal=args.traversal, queue_depth=args.queue_depth, waves_per_cu=args.waves_per_cu,
                     ramp_target=args.ramp_target, keep=args.keep, seed_budget=args.seed_budget)
    if args.bfs_max:
        ctx.set_option("bfs_max", args.bfs_max)
    if args.block_waves:
        ctx.set_option("block_waves", args.block_waves)
    if args.item_lanes:
        ctx.set_option("item_lanes", args.item_lanes)
    if args.node_slab:
        ctx.set_option("node_slab", args.node_slab)
    ids, pose_p, pose_q, pairs = build_pair(ctx, wl)
    dp, dq = device_poses(pose_p, pose_q)
    # the step chains refit and traversal on the context stream: the refit
    # returns once queued and the traversal call's completion covers both
    ctx.set_option("async_refit", 1)

    def step():
        ctx.refit_device(ids, dp.data_ptr(), dq.data_ptr())
        return ctx.tandem_traversal(pairs, narrow=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counters = None
    for _ in range(args.steps):
        counters = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    sched = ctx.traversal_stats()  # of the last timed step
    sched["block_waves"] = ctx.get_option("block_waves")
    # phase times and the dominant kernel (the persistent rounds): a second
    # timed pass of the same steps with HIP events on the context stream (phase
    # events, and one pair attached to each round's dispatch packet); events
    # cost a few us each, so the pass above runs without them
    ctx.set_option("kernel_timing", 2)
    step()
    torch.cuda.synchronize()
    trav_ms, narrow_ms, launch_ms, ptests, launches, round_ms = 0.0, 0.0, 0.0, 0, 0, None
    for _ in range(args.steps):
        step()
        tm, st = ctx.last_timing(), ctx.traversal_stats()
        trav_ms += tm["traversal_ms"]
        narrow_ms += tm["narrow_ms"]
        launch_ms += tm["per_launch_ms"] * st["persist_launches"]
        ptests += st["persist_tests"]
        launches += st["persist_launches"]
        r = tm["round_ms"][:st["persist_launches"]]
        if round_ms is None:
            round_ms, round_n = [], []
        for i, x in enumerate(r):  # each round's mean over the calls that launched it
            if i == len(round_ms):
                round_ms.append(0.0)
                round_n.append(0)
            round_ms[i] += x
            round_n[i] += 1
    torch.cuda.synchronize()
    ctx.set_option("kernel_timing", 0)
    trav_ms /= args.steps
    narrow_ms /= args.steps
    per_launch_ms = launch_ms / max(launches, 1)
    # a record-counting call (measurement, not timed): the distinct node records
    # the persistent rounds read; the set does not depend on the schedule
    ctx.set_option("count_records", 1)
    step()
    torch.cuda.synchronize()
    cst = ctx.traversal_stats()
    ctx.set_option("count_records", 0)
    # gather roofs of this node pool, loaded the way the expansion loads records:
    # random records within 16,384 slots (2 MB: L2-resident) and over the whole pool
    roof_l2 = ctx.gather_roof(window=16384)
    roof_pool = ctx.gather_roof(window=0)
    return dict(wl=wl, ctx=ctx, ids=ids, counters=counters, ms_per_step=ms_per_step, trav_ms=trav_ms,
                narrow_ms=narrow_ms, per_launch_ms=per_launch_ms, launches_per_step=launches / args.steps, sched=sched,
                ptests_per_launch=ptests / max(launches, 1), distinct_records=cst["distinct_records"],
                count_launches=cst["persist_launches"], records_loaded=cst["records_loaded"], roof_l2=roof_l2,
                roof_pool=roof_pool, round_ms=[x / n for x, n in zip(round_ms, round_n)] if round_ms else None)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy calibration);
    None when no summary for it is present."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return {"bytes": float(f[2]), "source": "profiles/pmc_traffic.txt"}
    return None


def batched_leg(args):
    import numpy as np
    import torch
    import kdop_amd as KA
    from kdop_amd import workloads as W
    wl = W.batched_pairs(npairs=args.batched_pairs, K=args.batched_k, seed=11)
    ctx = KA.Context(K=args.batched_k, traversal=args.traversal, queue_depth=args.queue_depth,
                     waves_per_cu=args.waves_per_cu)
    ids, pose_p, pose_q, pairs = build_pair(ctx, wl)
    dp, dq = device_poses(pose_p, pose_q)
    ctx.refit_device(ids, dp.data_ptr(), dq.data_ptr())
    reps = max(1, args.batched_reps)
    # the project's autotuner picks the traversal options on this very call
    tuned, tuned_ms = ctx.autotune(pairs, narrow=True, budget_ms=args.tune_ms) if args.tune_ms > 0 else (ctx.options(), None)
    # batched: every pair in one traversal call
    for _ in range(2):
        ctx.tandem_traversal(pairs, narrow=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        c = ctx.tandem_traversal(pairs, narrow=True)
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / reps
    # the same batched call under each traversal strategy with the tuned options
    # otherwise (the one-launch persistent design beside the level-synchronous one)
    strategy_ms = {}
    tuned_strategy = ctx.get_option("traversal")
    for name, mode in (("persistent", KA.KDOP_TRAVERSAL_PERSISTENT), ("levels", KA.KDOP_TRAVERSAL_LEVELS)):
        ctx.set_option("traversal", mode)
        for _ in range(2):
            ctx.tandem_traversal(pairs, narrow=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            cs = ctx.tandem_traversal(pairs, narrow=True)
        torch.cuda.synchronize()
        strategy_ms[name] = (time.perf_counter() - t0) / reps * 1e3
        assert cs["contacts"] == c["contacts"] and cs["tests"] == c["tests"], (name, cs, c)
    ctx.set_option("traversal", tuned_strategy)
    # unbatched: one call per body pair
    for i in range(min(8, pairs.shape[0])):
        ctx.tandem_traversal(pairs[i:i + 1], narrow=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot = 0
    for i in range(pairs.shape[0]):
        tot += ctx.tandem_traversal(pairs[i:i + 1], narrow=True)["contacts"]
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    assert tot == c["contacts"], (tot, c["contacts"])
    n = pairs.shape[0]
    # device contact sort of the batched call's contacts, in both report
    # orders (after a fresh traversal each time; the sort alone is timed),
    # beside a host lexsort of the same records
    ctx.tandem_traversal(pairs, narrow=True)
    con = ctx.contacts()
    t0 = time.perf_counter()
    np.lexsort((con["depth"], con["n"][:, 2], con["n"][:, 1], con["n"][:, 0], con["p"][:, 2], con["p"][:, 1], con["p"][:, 0],
                con["pair"]))
    t_host = time.perf_counter() - t0
    sort_ms = {}
    for order in ("sequential", "opencl"):
        tt = 0.0
        for _ in range(reps):
            ctx.tandem_traversal(pairs, narrow=True)
            t0 = time.perf_counter()
            ctx.sort_contacts(order)
            tt += time.perf_counter() - t0
        sort_ms[order] = tt / reps * 1e3
    # device PROX contact reduction of the same contacts (sort by the PROX key,
    # std::unique with the working precision), after a fresh traversal each time
    tt = 0.0
    for _ in range(reps):
        ctx.tandem_traversal(pairs, narrow=True)
        t0 = time.perf_counter()
        kept = ctx.reduce_contacts()
        tt += time.perf_counter() - t0
    reduce_ms = tt / reps * 1e3
    # the reduction's input as PROX would hold it, for the CPU baseline
    bp = np.asarray(pairs, np.uint32)
    prox_in = dict(bi=bp[con["pair"], 0], bj=bp[con["pair"], 1], p=con["p"].copy(), n=con["n"].copy(), d=con["depth"].copy())
    return dict(workload=wl.name, pairs=int(n), K=args.batched_k, tets_total=int(sum(wl.info["tets"])),
                separated_pairs=wl.info["separated"],
                pair_queries_per_s=n / t_batched, unbatched_pair_queries_per_s=n / t_single,
                batched_speedup=t_single / t_batched, batched_ms=t_batched * 1e3, unbatched_ms=t_single * 1e3,
                persistent_ms=strategy_ms["persistent"], levels_ms=strategy_ms["levels"],
                persistent_vs_levels=strategy_ms["persistent"] / strategy_ms["levels"],
                tests=c["tests"], contacts=c["contacts"], autotuned=tuned, autotune_best_ms=tuned_ms,
                contact_sort={"contacts": int(con.shape[0]), "device_ms": sort_ms,
                              "host_numpy_lexsort_opencl_order_ms": t_host * 1e3},
                contact_reduction={"contacts": int(con.shape[0]), "kept": int(kept), "device_ms": reduce_ms},
                _prox_in=prox_in)


def c1_leg(args):
    """The reference's own CPU-runnable case: two 10^3-cube boxes (5,000 tets
    each, K=8), B rotated by a seeded quaternion and moved by (0.05, 0.05,
    0.05).  Device: refit of both bodies + traversal + narrow phase per query,
    poses in device memory.  CPU: the oracle's sequential query (every branch
    row, one thread; the transform and refit are not timed on the CPU side)."""
    import numpy as np
    import torch
    import kdop_amd as KA
    from kdop_amd import workloads as W
    wl = W.single_pair(n=10, K=8, seed=1)
    ctx = KA.Context(K=wl.K)
    ids, pose_p, pose_q, pairs = build_pair(ctx, wl)
    dp, dq = device_poses(pose_p, pose_q)
    ctx.set_option("async_refit", 1)

    def step():
        ctx.refit_device(ids, dp.data_ptr(), dq.data_ptr())
        return ctx.tandem_traversal(pairs, narrow=True)

    for _ in range(20):
        step()
    torch.cuda.synchronize()
    n = 200
    t0 = time.perf_counter()
    for _ in range(n):
        c = step()
    torch.cuda.synchronize()
    dev_ms = (time.perf_counter() - t0) * 1e3 / n
    out = dict(workload=wl.name, tets_per_body=int(wl.meshes[0].nt), K=wl.K, device_ms_per_query=dev_ms,
               reference_tests=int(c["reference_tests"]), leaf_pairs=int(c["leaf_pairs"]), contacts=int(c["contacts"]))
    ctx.close()
    if not args.no_cpu:
        from oracle import oracle as O
        trees, meshes = [], []
        for m, pose in zip(wl.meshes, wl.poses):
            X, Y, Z = O.transform_f32(m.X, m.Y, m.Z, pose.p, pose.q)
            t = O.Tree(wl.K, 8000, m.tets, m.X, m.Y, m.Z)
            t.refit(X, Y, Z)
            trees.append(t)
            meshes.append((m.tets, X, Y, Z, m.surface))
        rows = trees[0].levels()[-1]
        best = None
        for _ in range(3):
            co, so = O.traverse_rows(trees[0], meshes[0], trees[1], meshes[1], 0, rows)
            best = so if best is None else min(best, so)
        out["cpu_baseline"] = dict(ms_per_query=best * 1e3, cores=1, kind="port",
                                   sample=f"oracle sequential tandem traversal + narrow phase, whole query ({co['tests']} tests), best of 3")
        # the row traversal makes every branch-row test; the reference adds the one root test
        out["same_counts"] = (co["tests"] + 1 == c["reference_tests"] and co["leaf_pairs"] == c["leaf_pairs"]
                              and co["contacts"] == c["contacts"])
        out["vs_cpu"] = best * 1e3 / dev_ms
    return out


def series_leg(args):
    """Refit + re-traverse time series (64 overlapping pairs of 20k-80k tets,
    K=18, 200 steps of seeded rigid motion): every step refits all bodies to
    that step's poses (device arrays) and traverses all pairs with the narrow
    phase, the loop a simulator runs.  The trees are built once and refitted,
    never rebuilt.  Reports the time per step over all 200 steps after one
    untimed pass of the first 10."""
    import numpy as np
    import torch
    import kdop_amd as KA
    from kdop_amd import workloads as W
    wl, P, Q = W.refit_series(npairs=64, K=18, steps=200)
    ctx = KA.Context(K=wl.K)
    ids = [ctx.make_tree(m) for m in wl.meshes]
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    Pd = torch.from_numpy(np.ascontiguousarray(P)).to("cuda")
    Qd = torch.from_numpy(np.ascontiguousarray(Q)).to("cuda")
    torch.cuda.synchronize()
    ctx.set_option("async_refit", 1)

    def step(s):
        ctx.refit_device(ids, Pd[s].data_ptr(), Qd[s].data_ptr())
        return ctx.tandem_traversal(pairs, narrow=True)

    for s in range(10):
        step(s)
    torch.cuda.synchronize()
    tests = contacts = surface = 0
    t0 = time.perf_counter()
    for s in range(P.shape[0]):
        c = step(s)
        tests += c["tests"]
        contacts += c["contacts"]
        surface += c["surface_pairs"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nsteps = P.shape[0]
    out = dict(workload=wl.name, pairs=len(wl.pairs), bodies=len(wl.meshes), K=wl.K,
               tets_total=int(sum(m.nt for m in wl.meshes)), steps=nsteps, ms_per_step=dt * 1e3 / nsteps,
               pair_queries_per_s=len(wl.pairs) * nsteps / dt, device_tests_per_s=tests / dt,
               tests_total=int(tests), surface_pairs_total=int(surface), contacts_total=int(contacts),
               note="per step: refit of all 128 bodies to that step's poses (device arrays) + tandem traversal + narrow "
                    "phase of all 64 pairs; trees built once, refitted every step")
    ctx.close()
    return out


def adjoint_leg(args):
    """The differentiable narrow phase on its configuration: 64 overlapping
    pairs of ~53k-tet jittered boxes, K=14.  Times the traversal that fixes
    the pair set and contacts, the adjoint of Phi (sum of contact depths per
    pair) for every body's pose, and one torch.autograd forward + backward
    (refit, traversal, adjoint).  The reference has no differentiable narrow
    phase, so there is no CPU path to time beside it."""
    import numpy as np
    import torch
    import kdop_amd as KA
    from kdop_amd import workloads as W
    from kdop_amd.autograd import penetration_measure
    wl = W.batched_pairs(npairs=64, K=14, seed=29, tmin=5e4, tmax=5e4, separated=0.0)
    ctx = KA.Context(K=14)
    ids, pose_p, pose_q, pairs = build_pair(ctx, wl)
    ids = np.asarray(ids, np.uint32)
    dp, dq = device_poses(pose_p, pose_q)
    ctx.refit_device(ids, dp.data_ptr(), dq.data_ptr())
    reps = 5
    for _ in range(2):
        c = ctx.tandem_traversal(pairs, narrow=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        c = ctx.tandem_traversal(pairs, narrow=True)
    torch.cuda.synchronize()
    t_trav = (time.perf_counter() - t0) / reps
    ctx.narrow_adjoint(ids)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        phi, gp, gq = ctx.narrow_adjoint(ids)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / reps
    p = torch.tensor(pose_p, device="cuda", requires_grad=True)
    q = torch.tensor(pose_q, device="cuda", requires_grad=True)
    penetration_measure(ctx, ids, pairs, p, q).sum().backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        p.grad = None
        q.grad = None
        penetration_measure(ctx, ids, pairs, p, q).sum().backward()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / reps
    return dict(workload="c5_64pairs_53k_k14", pairs=int(pairs.shape[0]), bodies=int(ids.shape[0]), tets_total=int(sum(wl.info["tets"])),
                contacts=int(c["contacts"]), narrow_pairs=int(c["narrow_pairs"]), traversal_ms=t_trav * 1e3, adjoint_ms=t_adj * 1e3,
                autograd_step_ms=t_step * 1e3,
                note="adjoint: Phi and dPhi/d(p, q) of all 128 bodies in FP64 from the frozen records; autograd step: "
                     "refit + traversal + narrow phase forward, adjoint backward")


def ray_set(n, seed=5):
    """Rays at the headline body A: origins on a sphere of radius 2 around its
    pose, aimed at points of its box (a few start inside or point away)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = 2.0 * u
    d = rng.uniform(-0.5, 0.5, size=(n, 3)) - o
    return o.astype(np.float32), d.astype(np.float32)


def raycast_leg(args, head):
    """kdop_raycast of many rays at body A of the headline pair (its refitted pose)."""
    import torch
    ctx, body = head["ctx"], head["ids"][0]
    pose = head["wl"].poses[0]
    o, d = ray_set(args.rays)
    o = o + pose.p[None, :]
    reps = 5
    # rays and results resident in device memory (is_device = 1)
    od, dd = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    ctx.raycast(body, od, dd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hits, lengths, flags = ctx.raycast(body, od, dd)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nhit = int(flags.sum())
    # host arrays in and out: the PCIe copies included
    ctx.raycast(body, o, d)
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, hflags = ctx.raycast(body, o, d)
    dth = (time.perf_counter() - t0) / reps
    assert int(hflags.sum()) == nhit
    return dict(rays=int(o.shape[0]), hits=nhit, ms=dt * 1e3, rays_per_s=o.shape[0] / dt, host_arrays_ms=dth * 1e3,
                body_tets=int(head["wl"].info["nt"]),
                note="rays and results in device memory; host_arrays_ms: host arrays in and out, copies included")


def sphere_set(n, seed=6):
    """Spheres around body A of the headline pair: centres in a box of half
    width 0.75 around its pose (near its surface and inside), radii 0.01-0.15."""
    import numpy as np
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.75, 0.75, size=(n, 3))
    r = rng.uniform(0.01, 0.15, size=(n, 1))
    return np.concatenate([c, r], 1).astype(np.float32)


def single_leg(args, head):
    """kdop_single_traversal of many spheres against body A of the headline pair."""
    import torch
    ctx, body = head["ctx"], head["ids"][0]
    sp = sphere_set(args.spheres)
    sp[:, :3] += head["wl"].poses[0].p[None, :]
    reps = 5
    # spheres and contact records resident in device memory (is_device = 1)
    spd = torch.from_numpy(sp).cuda()
    out = ctx.single_traversal(body, spd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = ctx.single_traversal(body, spd)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ncon = int(out.shape[0])
    # host arrays in and out: the PCIe copies included
    hout = ctx.single_traversal(body, sp)
    t0 = time.perf_counter()
    for _ in range(reps):
        hout = ctx.single_traversal(body, sp)
    dth = (time.perf_counter() - t0) / reps
    assert int(hout.shape[0]) == ncon
    return dict(spheres=int(sp.shape[0]), contacts=ncon, ms=dt * 1e3, spheres_per_s=sp.shape[0] / dt, host_arrays_ms=dth * 1e3,
                body_tets=int(head["wl"].info["nt"]),
                note="spheres and contact records in device memory; host_arrays_ms: host arrays in and out, copies included")


def reorder_leg(args):
    """Morton reorder of the headline mesh (40^3 box, 320k tets) on the device
    (host arrays in and out, copies included) against the library's host
    implementation of the same order."""
    import kdop_amd as KA
    X, Y, Z, T = KA.box_mesh_raw(args.n, size=1.0, jitter=0.1, seed=7)
    ctx = KA.Context(K=args.k)
    ctx.mesh_reorder(X, Y, Z, T)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        out = ctx.mesh_reorder(X, Y, Z, T)
    t_dev = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    host = KA.mesh_reorder(X, Y, Z, T)
    t_host = time.perf_counter() - t0
    same = all(a.tobytes() == b.tobytes() for a, b in zip(out, host))
    return dict(tets=int(T.shape[0]), verts=int(X.shape[0]), device_ms=t_dev * 1e3, host_ms=t_host * 1e3, identical=same)


def cpu_leg(args, head, batched=None):
    """Oracle (a restatement of the reference's sequential CPU traversal) on a
    bounded sample: the first R branch rows of body A against all of body B,
    one thread; rate scaled to the full query.  Also the oracle's raycast,
    single traversal and PROX contact reduction on the inputs of those legs."""
    import numpy as np
    from oracle import oracle as O
    wl = head["wl"]
    worlds, trees = [], []
    for m, pose in zip(wl.meshes, wl.poses):
        X, Y, Z = O.transform_f32(m.X, m.Y, m.Z, pose.p, pose.q)
        t = O.Tree(wl.K, 8000, m.tets, m.X, m.Y, m.Z)
        t.refit(X, Y, Z)
        worlds.append((X, Y, Z))
        trees.append(t)
    ma, mb = wl.meshes
    meshA = (ma.tets, *worlds[0], ma.surface)
    meshB = (mb.tets, *worlds[1], mb.surface)
    rows_total = trees[0].levels()[-1]
    c, s = O.traverse_rows(trees[0], meshA, trees[1], meshB, 0, 2)
    per_row = max(s / 2, 1e-6)
    rows = int(min(rows_total, max(2, args.cpu_seconds / per_row)))
    c, s = O.traverse_rows(trees[0], meshA, trees[1], meshB, 0, rows)
    tests_per_s = c["tests"] / s
    full_query_s = s * rows_total / rows
    one = dict(value=tests_per_s, unit="node-pair tests/s", cores=1, kind="port",
               sample=f"oracle sequential tandem traversal + narrow phase, branch rows 0..{rows} of {rows_total} "
                      f"of body A vs all {trees[1].levels()[-1]} branches of body B ({c['tests']} tests in {s:.2f} s)",
               est_full_query_s=full_query_s, sample_seconds=s)
    # kdop::raycast as the reference runs it (every branch root per ray), one thread, on a sample of the rays
    ray = None
    if head.get("raycast"):
        o, d = ray_set(head["raycast"]["rays"])
        o = o + wl.poses[0].p[None, :]
        nr = 8
        import time as _t
        t0 = _t.perf_counter()
        O.raycast(trees[0], meshA, o[:nr], d[:nr])
        per = (_t.perf_counter() - t0) / nr
        nr = int(min(o.shape[0], max(nr, args.cpu_seconds / 3 / max(per, 1e-6))))
        t0 = _t.perf_counter()
        O.raycast(trees[0], meshA, o[:nr], d[:nr])
        sr = _t.perf_counter() - t0
        ray = dict(value=nr / sr, unit="rays/s", cores=1, kind="port", sample=f"oracle kdop::raycast, first {nr} rays of the set ({sr:.2f} s)")
    sph = None
    if head.get("single"):
        sp = sphere_set(head["single"]["spheres"])
        sp[:, :3] += wl.poses[0].p[None, :]
        import time as _t
        nr = 8
        t0 = _t.perf_counter()
        O.single_traversal(trees[0], meshA, sp[:nr])
        per = (_t.perf_counter() - t0) / nr
        nr = int(min(sp.shape[0], max(nr, args.cpu_seconds / 3 / max(per, 1e-6))))
        t0 = _t.perf_counter()
        O.single_traversal(trees[0], meshA, sp[:nr])
        ss = _t.perf_counter() - t0
        sph = dict(value=nr / ss, unit="spheres/s", cores=1, kind="port",
                   sample=f"oracle kdop::single_traversal, first {nr} spheres of the set ({ss:.2f} s)")
    # PROX contact reduction (std::sort + std::unique) of the batched call's contacts, one thread
    red = None
    if batched and batched.get("_prox_in") is not None:
        import time as _t
        pi = batched["_prox_in"]
        t0 = _t.perf_counter()
        kept = O.prox_reduce(pi["bi"], pi["bj"], pi["p"], pi["n"], pi["d"])
        sr = _t.perf_counter() - t0
        red = dict(ms=sr * 1e3, kept=int(kept.shape[0]), cores=1, kind="port",
                   sample=f"oracle std::stable_sort + std::unique of the batched call's {pi['d'].shape[0]} contacts")
    # the same query on every host core this process may run on (rows taken one
    # at a time), whole query
    host = host_cpu_info()
    cores = host["affinity_cpus"]
    cm, sm = O.traverse_rows_mt(trees[0], meshA, trees[1], meshB, 0, rows_total, cores)
    allc = dict(value=cm["tests"] / sm, unit="node-pair tests/s", cores=cores, kind="port",
                sample=f"oracle tandem traversal + narrow phase, all {rows_total} branch rows of body A on {cores} threads "
                       f"({cm['tests']} tests in {sm:.2f} s)", full_query_s=sm, host=host)
    one["host"] = host
    return one, allc, ray, sph, red


def host_cpu_info():
    """The host CPUs: nproc, the CPUs this process may run on, the cgroup CPU
    quota (cpu.max) and the model name (lscpu)."""
    import subprocess
    info = dict(nproc=os.cpu_count(), affinity_cpus=len(os.sched_getaffinity(0)))
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        info["cgroup_cpu_quota"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        info["cgroup_cpu_quota"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def contact_fp64_error(head):
    """Every contact of the headline query against the FP64 restatement of its
    tet pair with the float32 run's discrete choices (face, candidates, extreme
    points) frozen: the largest error of points, normal and depth relative to
    max(|x|, 1).  Also checks that the device's float32 contacts are the
    oracle's, bit for bit."""
    import numpy as np
    from oracle import oracle as O
    wl, ctx = head["wl"], head["ctx"]
    con = ctx.contacts()
    ma, mb = wl.meshes
    PA = np.stack(O.transform_f32(ma.X, ma.Y, ma.Z, wl.poses[0].p, wl.poses[0].q), 1)
    PB = np.stack(O.transform_f32(mb.X, mb.Y, mb.Z, wl.poses[1].p, wl.poses[1].q), 1)
    key = np.lexsort((con["ordinal"], con["tet_b"], con["tet_a"]))
    con = con[key]
    worst, pairs, i = 0.0, 0, 0
    while i < con.shape[0]:
        j = i
        while j < con.shape[0] and con["tet_a"][j] == con["tet_a"][i] and con["tet_b"][j] == con["tet_b"][i]:
            j += 1
        ta, tb = int(con["tet_a"][i]), int(con["tet_b"][i])
        r = O.contacts_pair_codes(PA[ma.tets[ta]], PB[mb.tets[tb]], ma.surface[ta], mb.surface[tb])
        assert r["pts"].shape[0] == j - i and r["pts"].tobytes() == con["p"][i:j].tobytes(), (ta, tb)
        n64, d64, p64 = O.frozen_pair_f64(PA[ma.tets[ta]], PB[mb.tets[tb]], r["face"], r["codes"], r["uniq"], r["imin"], r["imax"])
        for got, want in ((con["p"][i:j], p64), (con["n"][i:j], np.broadcast_to(n64, (j - i, 3))),
                          (con["depth"][i:j], np.full(j - i, d64))):
            err = np.abs(got.astype(np.float64) - want) / np.maximum(np.abs(want), 1.0)
            worst = max(worst, float(err.max()))
        pairs += 1
        i = j
    return dict(contacts=int(con.shape[0]), pairs=pairs, max_rel_err=worst,
                note="device float32 contacts vs the FP64 restatement with the float32 choices frozen; "
                     "relative to max(|x|, 1)")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--k", type=int, default=26)
    ap.add_argument("--n", type=int, default=40, help="cubes per box edge (5 n^3 tets per body)")
    ap.add_argument("--batched-pairs", type=int, default=512)
    ap.add_argument("--batched-k", type=int, default=8)
    ap.add_argument("--batched-reps", type=int, default=20)
    ap.add_argument("--tune-ms", type=float, default=3000.0, help="autotuner budget for the batched leg (0: off)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traversal", type=int, default=0, help="0 auto, 1 level-synchronous, 2 persistent")
    ap.add_argument("--queue-depth", type=int, default=0, help="persistent kernel: LDS stack entries per wave")
    ap.add_argument("--waves-per-cu", type=int, default=0, help="persistent kernel: waves per CU")
    ap.add_argument("--ramp-target", type=int, default=0, help="persistent kernel: frontier size that ends the ramp")
    ap.add_argument("--keep", type=int, default=0, help="persistent kernel: stack kept once the frontier is gone")
    ap.add_argument("--seed-budget", type=int, default=0, help="chunk-root pairs tested to seed a call")
    ap.add_argument("--bfs-max", type=int, default=0, help="persistent strategy: most breadth-first levels first")
    ap.add_argument("--block-waves", type=int, default=0, help="persistent kernel: waves sharing one stack (1, 4, 8)")
    ap.add_argument("--item-lanes", type=int, default=0, help="persistent kernel: lanes per work item (1 or 4)")
    ap.add_argument("--node-slab", type=int, default=0, help="persistent kernel: 1 stages node records in LDS node slabs")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-reorder", action="store_true", help="skip the device Morton reorder leg")
    ap.add_argument("--rays", type=int, default=65536, help="rays of the raycast leg (0: skip)")
    ap.add_argument("--spheres", type=int, default=65536, help="spheres of the single-traversal leg (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable narrow phase leg")
    ap.add_argument("--no-series", action="store_true", help="skip the refit + re-traverse time series leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a device-to-device copy of MB megabytes (reference for rocprof byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the kDOP pipeline here is a single-device build "
              "(body pairs would shard across GPUs, see the design notes); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401  (shared HIP runtime, device pose buffers, synchronisation)
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log(f"[bench] headline: box{args.n} pair, K={args.k}, steps={args.steps}, warmup={args.warmup}")
    head = headline_leg(args)
    cnt = head["counters"]
    log(f"[bench] {head['ms_per_step']:.3f} ms/step, traversal {head['trav_ms']:.3f} ms, narrow {head['narrow_ms']:.3f} ms")
    batched = None
    if not args.no_batched:
        log(f"[bench] batched leg: {args.batched_pairs} pairs")
        batched = batched_leg(args)
        log(f"[bench] batched {batched['pair_queries_per_s']:.1f} pair-queries/s, speedup {batched['batched_speedup']:.1f}x")
    if args.rays:
        head["raycast"] = raycast_leg(args, head)
        log(f"[bench] raycast {head['raycast']['rays']} rays: {head['raycast']['ms']:.2f} ms")
    if args.spheres:
        head["single"] = single_leg(args, head)
        log(f"[bench] single traversal {head['single']['spheres']} spheres: {head['single']['ms']:.2f} ms")
    c1 = c1_leg(args)
    log(f"[bench] C1 (5k-tet pair, K=8): {c1['device_ms_per_query']:.3f} ms per query on the device"
        + (f", {c1['cpu_baseline']['ms_per_query']:.2f} ms on one core" if c1.get('cpu_baseline') else ""))
    series = None
    if not args.no_series:
        series = series_leg(args)
        log(f"[bench] refit series (64 pairs, K=18, 200 steps): {series['ms_per_step']:.3f} ms per step, "
            f"{series['pair_queries_per_s']:.0f} pair-queries/s")
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args)
        log(f"[bench] adjoint (64 pairs, K=14): traversal {adjoint['traversal_ms']:.2f} ms, adjoint {adjoint['adjoint_ms']:.2f} ms, "
            f"autograd step {adjoint['autograd_step_ms']:.2f} ms")
    reorder = None
    if not args.no_reorder:
        reorder = reorder_leg(args)
        log(f"[bench] mesh reorder {reorder['tets']} tets: device {reorder['device_ms']:.2f} ms, host {reorder['host_ms']:.2f} ms")
    cpu = cpu_all = None
    if not args.no_cpu:
        log("[bench] cpu baseline sample")
        cpu, cpu_all, cpu_ray, cpu_sph, cpu_red = cpu_leg(args, head, batched)
        if cpu_red and batched:
            cr = batched["contact_reduction"]
            cr["cpu_baseline"] = cpu_red
            cr["vs_cpu"] = cpu_red["ms"] / cr["device_ms"]
            cr["identical_count"] = cpu_red["kept"] == cr["kept"]
        if cpu_sph and head.get("single"):
            head["single"]["cpu_baseline"] = cpu_sph
            head["single"]["vs_cpu"] = head["single"]["spheres_per_s"] / cpu_sph["value"]
        if cpu_ray and head.get("raycast"):
            head["raycast"]["cpu_baseline"] = cpu_ray
            head["raycast"]["vs_cpu"] = head["raycast"]["rays_per_s"] / cpu_ray["value"]
        log(f"[bench] cpu {cpu['value']:.3e} tests/s, est. full query {cpu['est_full_query_s']:.1f} s; "
            f"{cpu_all['cores']} cores: {cpu_all['full_query_s']:.2f} s")
    if batched:
        batched.pop("_prox_in", None)
    fp64 = None
    if not args.no_cpu:
        log("[bench] contacts vs the FP64 restatement")
        fp64 = contact_fp64_error(head)
    steps_per_s = 1e3 / head["ms_per_step"]
    sc = head["sched"]
    bw, ls = sc["block_waves"], sc["queue_depth"]
    if sc["item_lanes"] == 4:
        kernel = f"k_persist_quad<{args.k},{bw},{(ls * bw).bit_length() - 1}>"
    else:
        kernel = f"k_persist_blk<{args.k},512,{bw}>" if bw > 1 else f"k_persist<{args.k},{ls}>"
    traffic = measured_traffic(kernel)
    launch_ms = head["per_launch_ms"]
    rec_line = 4 * rec_stride(args.k)                    # bytes of one node record (one cache line at K >= 18)
    rec_used = 16 * ((args.k + 2 + 3) // 4)              # the words a record load reads
    cl = max(head["count_launches"], 1)
    compulsory = head["distinct_records"] / cl * rec_line  # each distinct record read from memory once
    gathered = head["records_loaded"] / cl * rec_used      # the record bytes the expansions load
    gbs = lambda b: b / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": cnt["tests"] * steps_per_s,
        "unit": "node-pair tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"box{args.n}_jitter_pair_k{args.k}", "tets_per_body": head["wl"].info["nt"],
                   "K": args.k, "chunk_bytes": 8000, "aabb_overlap": round(head["wl"].info["aabb_overlap"], 3),
                   "step": "refit both bodies + tandem traversal + narrow phase"},
        "roofline": {"bound": "hbm", "achieved": gbs(compulsory), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs(compulsory) / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_gbs": gbs(traffic["bytes"]) if traffic else None,
                     "traffic_frac": gbs(traffic["bytes"]) / HBM_PEAK_GBS if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": kernel, "launch_ms": launch_ms, "round_ms": head["round_ms"],
                     "launches_per_step": head["launches_per_step"],
                     "algorithmic_bytes_per_launch": compulsory,
                     "algorithmic": f"distinct node records the launch reads ({head['distinct_records'] / cl:.0f}, "
                                    f"counted on the device) x {rec_line} B: every one must come from memory at least once"},
        "gather": {"bytes_per_launch": gathered, "achieved": gbs(gathered), "unit": "GB/s",
                   "roof_l2": head["roof_l2"], "frac_l2": gbs(gathered) / head["roof_l2"] if head["roof_l2"] else None,
                   "roof_pool": head["roof_pool"], "frac_pool": gbs(gathered) / head["roof_pool"] if head["roof_pool"] else None,
                   "records_per_launch": head["records_loaded"] / cl, "bytes_per_record": rec_used,
                   "tests_per_launch": head["ptests_per_launch"],
                   "survey_bytes_per_test": 2 * (4 * args.k + 8),
                   "note": "record bytes the expansions load (one load per record and item) per launch time, against the "
                           "rate of the same loads of random records within 2 MB (roof_l2) and over the whole node pool "
                           "(roof_pool), measured on this GPU by kdop_gather_roof"},
        "contacts_fp64": fp64,
        "counters": cnt,
        "reference_tests_per_s": cnt["reference_tests"] * steps_per_s,
        "traversal_ms": head["trav_ms"],
        "schedule": head["sched"],
        "narrow_ms": head["narrow_ms"],
        "batched": batched,
        "mesh_reorder": reorder,
        "raycast": head.get("raycast"),
        "single_traversal": head.get("single"),
        "c1_reference_case": c1,
        "refit_series": series,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    if cpu:
        out["vs_cpu_query_speedup"] = cpu["est_full_query_s"] * 1e3 / head["ms_per_step"]
        out["vs_cpu_all_cores_query_speedup"] = cpu_all["full_query_s"] * 1e3 / head["ms_per_step"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
