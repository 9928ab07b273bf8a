"""Shared helpers of the test suite: run one pair workload through the oracle
and through the HIP library and return both results side by side."""
import numpy as np

import kdop_amd as K
from oracle import oracle as O


def world_coords(mesh, pose):
    """World vertex coordinates by the oracle's transform (reference op order)."""
    return O.transform_f32(mesh.X, mesh.Y, mesh.Z, pose.p, pose.q)


def oracle_run(wl, narrow=True, threads=1, chunk_bytes=8000):
    """Oracle traversal of every body pair of a workload.  Returns a list of
    (counters, p1, p2, contacts) per pair and the oracle trees per mesh.
    chunk_bytes: make_tree's mem_bytes, one value or one per mesh."""
    trees, worlds = [], []
    per_mesh = chunk_bytes if isinstance(chunk_bytes, (list, tuple)) else [chunk_bytes] * len(wl.meshes)
    for m, pose, mb in zip(wl.meshes, wl.poses, per_mesh):
        X, Y, Z = world_coords(m, pose)
        t = O.Tree(wl.K, mb, m.tets, m.X, m.Y, m.Z)  # make_tree on material coordinates
        t.refit(X, Y, Z)                                 # update_kdop_bvh
        trees.append(t)
        worlds.append((X, Y, Z))
    out = []
    for a, b in wl.pairs:
        ma, mb = wl.meshes[a], wl.meshes[b]
        res = O.traverse(trees[a], (ma.tets, *worlds[a], ma.surface), trees[b], (mb.tets, *worlds[b], mb.surface),
                         narrow=narrow, threads=threads)
        out.append(res[:4])
    return out, trees, worlds


def device_run(wl, narrow=True, keep_all=True, ctx=None, ids=None, **ctx_kw):
    """Build (unless the body ids are given), refit and traverse a workload on the device."""
    own = ctx is None
    if own:
        ctx = K.Context(K=wl.K, **ctx_kw)
    if ids is None:
        ids = [ctx.make_tree(m) for m in wl.meshes]
    ctx.refit(ids, np.stack([p.p for p in wl.poses]), np.stack([p.q for p in wl.poses]))
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    cnt = ctx.tandem_traversal(pairs, narrow=narrow, keep_all_pairs=keep_all)
    res = dict(counters=cnt, p2=ctx.pairs(1), ids=ids, ctx=ctx)
    if keep_all:
        res["p1"] = ctx.pairs(0)
    if narrow:
        res["contacts"] = ctx.contacts()
    return res


def sort_pairs(p):
    """Sort an (n, k) uint32 pair array lexicographically (last column fastest)."""
    if p.shape[0] == 0:
        return p
    keys = tuple(p[:, i] for i in range(p.shape[1] - 1, -1, -1))
    return p[np.lexsort(keys)]


def oracle_contacts_with_ordinals(con):
    """Attach per-(ta, tb) ordinals to the oracle's contacts (report order)."""
    ords = np.zeros(con.shape[0], np.uint32)
    last = None
    k = 0
    for i, c in enumerate(con):
        key = (int(c["ta"]), int(c["tb"]))
        k = k + 1 if key == last else 0
        last = key
        ords[i] = k
    return ords


def narrow_records(wl, pair_index=0):
    """Float32 narrow records of one body pair of a workload, via the oracle."""
    orc, trees, worlds = oracle_run(wl)
    a, b = wl.pairs[pair_index]
    ma, mb = wl.meshes[a], wl.meshes[b]
    PA = np.stack(worlds[a], 1)
    PB = np.stack(worlds[b], 1)
    recs = []
    for ta_i, tb_i in orc[pair_index][2]:
        r = O.contacts_pair_codes(PA[ma.tets[ta_i]], PB[mb.tets[tb_i]], ma.surface[ta_i], mb.surface[tb_i])
        if r["pts"].shape[0]:
            recs.append((int(ta_i), int(tb_i), r))
    return recs, PA, PB, ma, mb
