"""Pins of the CPU oracle (oracle/kdop_oracle.cpp) by facts that do not come
from the oracle itself: closed forms, brute force, an LP feasibility check,
analytic contact configurations, and float64 restatements.  The reference
cannot be compiled here (its TINY headers need Boost), so these are what fix
the oracle; the HIP library is then checked against the oracle on the GPU."""
import math

import numpy as np
import pytest

import kdop_amd as K
from oracle import oracle as O
from helpers import narrow_records

F32 = np.float32


# ------------------------------------------------------------ direction tables
def test_direction_tables_closed_form():
    s3 = F32(1) / np.sqrt(F32(3))  # unit(): v * (1/sqrt(v.v)), correctly rounded float ops
    s2 = F32(1) / np.sqrt(F32(2))
    axes = np.eye(3, dtype=F32)
    diag = np.array([[1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1]], F32) * s3
    edge = np.array([[1, 1, 0], [1, -1, 0], [1, 0, 1], [0, 1, 1], [-1, 0, 1], [0, -1, 1]], F32) * s2
    want = {6: axes, 8: diag, 12: edge, 14: np.vstack([axes, diag]), 18: np.vstack([axes, edge]),
            20: np.vstack([diag, edge]), 26: np.vstack([axes, diag, edge])}
    for k, w in want.items():
        got = O.dir_table(k)
        assert got.shape == (k // 2, 3)
        np.testing.assert_array_equal(got.view(np.uint32), w.astype(F32).view(np.uint32))


# ------------------------------------------------------------------ make_dop
def _proj32(d, p):
    """((0 + d0 p0) + d1 p1) + d2 p2 in float32, one rounding per operation."""
    acc = F32(0)
    for j in range(3):
        acc = F32(acc + F32(d[j] * p[j]))
    return acc


def test_leaf_dop_k6_is_exact_aabb():
    rng = np.random.default_rng(0)
    for _ in range(200):
        pts = rng.normal(size=(4, 3)).astype(F32)
        dop = O.make_dop(6, pts)
        lo, hi = pts.min(0), pts.max(0)
        np.testing.assert_array_equal(dop[0::2], lo)
        np.testing.assert_array_equal(dop[1::2], hi)


@pytest.mark.parametrize("k", [8, 14, 18, 26])
def test_leaf_dop_contains_and_touches_vertices(k):
    rng = np.random.default_rng(k)
    D = O.dir_table(k)
    for _ in range(100):
        pts = rng.normal(size=(4, 3)).astype(F32)
        dop = O.make_dop(k, pts)
        exact = pts.astype(np.float64) @ D.astype(np.float64).T  # (4, N)
        lo, hi = dop[0::2].astype(np.float64), dop[1::2].astype(np.float64)
        tol = 4 * np.finfo(F32).eps * np.abs(pts).max() * 3
        assert np.all(lo <= exact.min(0) + tol) and np.all(hi >= exact.max(0) - tol)
        assert np.all(np.abs(lo - exact.min(0)) <= tol) and np.all(np.abs(hi - exact.max(0)) <= tol)
        # bits: every bound is one of the float32 projections with the reference's left-to-right sum
        for s in range(k // 2):
            proj = [_proj32(D[s], p) for p in pts]
            assert dop[2 * s] == min(proj) and dop[2 * s + 1] == max(proj)


# ---------------------------------------------------------------- tree shape
def _shape_closed_form(k, mem, m):
    node = 4 * k + 24
    n_max = math.floor(float(F32(mem // node) * F32(0.5)))
    lp = 1 << 30
    while lp > n_max and lp > 0:
        lp >>= 1
    up = 1
    while up < 2 * m:
        up <<= 1
    n_perfect = min(lp, up) - 1
    L = int(F32(n_perfect + 1) * F32(0.5))
    C = int(math.ceil(float(F32(F32(m) / F32(L)))))
    H = 1 if C <= 1 else int(max(math.ceil(math.log(C) / math.log(L)), 1.0))
    return dict(node_bytes=node, n_perfect=n_perfect, L=L, C=C, H=H)


@pytest.mark.parametrize("k", [6, 8, 12, 14, 18, 20, 26])
@pytest.mark.parametrize("m", [2, 3, 31, 32, 33, 1000, 5000, 8640, 50000, 320000])
def test_tree_shape_formula(k, m):
    want = _shape_closed_form(k, 8000, m)
    assert O.tree_shape(k, 8000, m) == want
    assert K.tree_shape(k, 8000, m) == want


def _median_tree(n):
    """Median split of n elements (kdop_make_tree.h:85): nested tuples, leaves = None."""
    if n == 1:
        return None
    left = (n - 1) // 2 + 1
    return (_median_tree(left), _median_tree(n - left))


def _chunk_sizes(m, L):
    C = -(-m // L)
    return [L] * (C - 1) + [m - (C - 1) * L]


def test_tree_levels_match_closed_form():
    for k, m in ((8, 5000), (26, 4000), (14, 20000), (6, 70000)):
        shp = _shape_closed_form(k, 8000, m)
        rng = np.random.default_rng(m)
        X, Y, Z = (rng.random(m + 3).astype(F32) for _ in range(3))
        tets = np.stack([np.arange(m), np.arange(m) + 1, np.arange(m) + 2, np.arange(m) + 3], 1).astype(np.uint32)
        t = O.Tree(k, 8000, tets, X, Y, Z)
        lv = t.levels()
        assert len(lv) == shp["H"] and lv[-1] == shp["C"]
        c = shp["C"]
        for h in range(shp["H"] - 2, -1, -1):
            c = int(math.ceil(float(F32(F32(c) / F32(shp["L"])))))
            assert lv[h] == c


# ------------------------------------------------------------ refit / unions
def test_refit_unions_exact_and_root_is_brute_force():
    m = K.box_mesh(6, jitter=0.2, seed=3)
    q = K.rng_quaternions(5, 0, 1)[0]
    X, Y, Z = O.transform_f32(m.X, m.Y, m.Z, np.array([0.1, -0.2, 0.3], F32), q)
    for k in (8, 26):
        t = O.Tree(k, 8000, m.tets, m.X, m.Y, m.Z)
        t.refit(X, Y, Z)
        ex = t.export()
        info, vol = ex["info"], ex["vol"]
        index = {(int(a), int(b), int(c)): i for i, (a, b, c, *_r) in enumerate(info)}
        nlev = len(t.levels())
        for i, (lv, ch, sl, leaf, st, en) in enumerate(info):
            if leaf:
                if lv == nlev - 1:  # tet leaf: every vertex inside (exact float32 projections)
                    D = O.dir_table(k)
                    pts = np.stack([X, Y, Z], 1)[m.tets[st]]
                    for s in range(k // 2):
                        pr = [_proj32(D[s], p) for p in pts]
                        assert vol[i, 2 * s] == min(pr) and vol[i, 2 * s + 1] == max(pr)
                else:  # super leaf: copy of the child chunk's root
                    np.testing.assert_array_equal(vol[i], vol[index[(int(lv) + 1, int(st), 0)]])
                continue
            a, b = vol[index[(int(lv), int(ch), int(st))]], vol[index[(int(lv), int(ch), int(en))]]
            np.testing.assert_array_equal(vol[i, 0::2], np.minimum(a[0::2], b[0::2]))
            np.testing.assert_array_equal(vol[i, 1::2], np.maximum(a[1::2], b[1::2]))
        # root = min/max over every vertex used by a tet
        used = np.unique(m.tets)
        D = O.dir_table(k)
        P = np.stack([X, Y, Z], 1)[used]
        root = t.root()
        for s in range(k // 2):
            pr = np.array([_proj32(D[s], p) for p in P], F32)
            assert root[2 * s] == pr.min() and root[2 * s + 1] == pr.max()


# ---------------------------------------------------------------- traversal
def _mesh_tuple(mesh, X=None, Y=None, Z=None):
    return (mesh.tets, mesh.X if X is None else X, mesh.Y if Y is None else Y, mesh.Z if Z is None else Z, mesh.surface)


def _numpy_leaf_pairs(k, ta, tb):
    """Independent brute force: leaf DOPs from the exported leaves, all pairs, strict overlap."""
    def leaves(t):
        ex = t.export()
        nlev = len(t.levels())
        sel = (ex["info"][:, 0] == nlev - 1) & (ex["info"][:, 3] == 1)
        return ex["info"][sel, 4].astype(np.uint32), ex["vol"][sel]
    ia, va = leaves(ta)
    ib, vb = leaves(tb)
    lo_a, hi_a = va[:, None, 0::2], va[:, None, 1::2]
    lo_b, hi_b = vb[None, :, 0::2], vb[None, :, 1::2]
    sep = ((hi_a < lo_b) | (hi_b < lo_a)).any(-1)
    i, j = np.nonzero(~sep)
    out = np.stack([ia[i], ib[j]], 1)
    return out[np.lexsort((out[:, 1], out[:, 0]))]


@pytest.mark.parametrize("k", [6, 8, 18, 26])
def test_pair_set_equals_brute_force(k):
    ma = K.box_mesh(6, jitter=0.15, seed=1)
    mb = K.box_mesh(5, jitter=0.15, seed=2)
    q = K.rng_quaternions(9, 0, 1)[0]
    Xb, Yb, Zb = O.transform_f32(mb.X, mb.Y, mb.Z, np.array([0.3, 0.1, -0.2], F32), q)
    ta = O.Tree(k, 8000, ma.tets, ma.X, ma.Y, ma.Z)
    tb = O.Tree(k, 8000, mb.tets, mb.X, mb.Y, mb.Z)
    tb.refit(Xb, Yb, Zb)
    c, p1, p2, con, _ = O.traverse(ta, _mesh_tuple(ma), tb, _mesh_tuple(mb, Xb, Yb, Zb))
    got = p1[np.lexsort((p1[:, 1], p1[:, 0]))]
    want = _numpy_leaf_pairs(k, ta, tb)
    np.testing.assert_array_equal(got, want)
    bp, nb = O.brute_pairs(ta, tb, 10 ** 7)
    assert nb == got.shape[0]
    # P2 = P1 restricted to tets that both have a surface face (kdop_tandem_traversal.h:61-88)
    keep = (ma.surface[got[:, 0]] != 0) & (mb.surface[got[:, 1]] != 0)
    p2s = p2[np.lexsort((p2[:, 1], p2[:, 0]))]
    np.testing.assert_array_equal(p2s, got[keep])


def test_separated_bodies_one_test():
    m = K.box_mesh(3)
    t = O.Tree(8, 8000, m.tets, m.X, m.Y, m.Z)
    X = m.X + F32(5.0)
    t2 = O.Tree(8, 8000, m.tets, X, m.Y, m.Z)
    c, p1, p2, con, _ = O.traverse(t, _mesh_tuple(m), t2, _mesh_tuple(m, X))
    assert c["tests"] == 1 and c["leaf_pairs"] == 0


def _two_cluster_mesh(n, gap):
    """Two n^3 boxes side by side in one mesh (tets of box 1 first)."""
    X, Y, Z, T = K.box_mesh_raw(n)
    X2 = X + F32(gap)
    Xc = np.concatenate([X, X2]).astype(F32)
    Yc = np.concatenate([Y, Y]).astype(F32)
    Zc = np.concatenate([Z, Z]).astype(F32)
    Tc = np.concatenate([T, T + X.shape[0]]).astype(np.uint32)
    return K.Mesh(Xc, Yc, Zc, Tc, K.compute_surface_map(Xc.shape[0], Tc))


def test_only_roots_overlap_gives_one_plus_branch_pairs():
    """Body A = two boxes 3 apart, body B = one box in the gap (K=6 boxes):
    the root boxes overlap, no branch root box does: 1 + C_A*C_B tests."""
    ma = _two_cluster_mesh(4, 3.0)   # 2 x 320 tets = 20 chunks of 32, chunk-aligned clusters
    X, Y, Z, T = K.box_mesh_raw(4)
    mb = K.Mesh(X + F32(1.5), Y, Z, T, K.compute_surface_map(X.shape[0], T))
    ta = O.Tree(6, 8000, ma.tets, ma.X, ma.Y, ma.Z)
    tb = O.Tree(6, 8000, mb.tets, mb.X, mb.Y, mb.Z)
    ca, cb = ta.levels()[-1], tb.levels()[-1]
    assert (ca, cb) == (20, 10)
    c, p1, p2, con, _ = O.traverse(ta, _mesh_tuple(ma), tb, _mesh_tuple(mb))
    assert c["tests"] == 1 + ca * cb and c["leaf_pairs"] == 0 and c["branch_overlaps"] == 0


def _star_mesh(n):
    """Every tet = (origin, surface triangle of an n^3 box): all leaf DOPs contain 0."""
    mesh = K.box_mesh(n)
    tris = []
    opp = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    for t, s in zip(mesh.tets, mesh.surface):
        for f in range(4):
            if (s >> f) & 1:
                tris.append([t[i] for i in opp[f]])
    tris = np.array(tris, np.uint32) + 1
    X = np.concatenate([[0.0], mesh.X]).astype(F32)
    Y = np.concatenate([[0.0], mesh.Y]).astype(F32)
    Z = np.concatenate([[0.0], mesh.Z]).astype(F32)
    T = np.concatenate([np.zeros((tris.shape[0], 1), np.uint32), tris], 1)
    return K.Mesh(X, Y, Z, T, np.full(T.shape[0], 1, np.uint8))


def _bvtt_count(ta, tb):
    """Node pairs the reference rule visits when every pair overlaps."""
    if ta is None and tb is None:
        return 1
    if ta is not None and tb is not None:
        return 1 + sum(_bvtt_count(x, y) for x in ta for y in tb)
    if ta is not None:
        return 1 + sum(_bvtt_count(x, tb) for x in ta)
    return 1 + sum(_bvtt_count(ta, y) for y in tb)


@pytest.mark.parametrize("k", [6, 8, 26])
def test_all_pairs_overlap_closed_form_count(k):
    ma, mb = _star_mesh(2), _star_mesh(3)   # 48 and 108 tets, all containing the origin
    ta = O.Tree(k, 8000, ma.tets, ma.X, ma.Y, ma.Z)
    tb = O.Tree(k, 8000, mb.tets, mb.X, mb.Y, mb.Z)
    L = _shape_closed_form(k, 8000, ma.nt)["L"]
    Lb = _shape_closed_form(k, 8000, mb.nt)["L"]
    sa = [_median_tree(s) for s in _chunk_sizes(ma.nt, L)]
    sb = [_median_tree(s) for s in _chunk_sizes(mb.nt, Lb)]
    want = 1 + sum(_bvtt_count(x, y) for x in sa for y in sb)
    c, p1, p2, con, _ = O.traverse(ta, _mesh_tuple(ma), tb, _mesh_tuple(mb), narrow=False)
    assert c["tests"] == want
    assert c["leaf_pairs"] == ma.nt * mb.nt
    if L == 32:  # perfect 32-leaf chunks give (4^6 - 1) / 3 pairs each
        assert _bvtt_count(_median_tree(32), _median_tree(32)) == (4 ** 6 - 1) // 3


# ------------------------------------------------- sequential visit order
def _leaf_tree(n, first):
    """_median_tree over tets first .. first + n - 1, leaves = tet ids."""
    if n == 1:
        return first
    left = (n - 1) // 2 + 1
    return (_leaf_tree(left, first), _leaf_tree(n - left, first + left))


def _branch_trees(nt, L):
    out, first = [], 0
    for s in _chunk_sizes(nt, L):
        out.append(_leaf_tree(s, first))
        first += s
    return out


def _aabb(tree, box):
    """K = 6 volume of a subtree: leaves from box (nt, 6) = per-tet (lo, hi) per axis."""
    if isinstance(tree, int):
        return box[tree]
    a, b = _aabb(tree[0], box), _aabb(tree[1], box)
    return np.array([min(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), max(a[3], b[3]), min(a[4], b[4]), max(a[5], b[5])])


def _visit_order(sa, sb, box_a, box_b):
    """Leaf pairs in the order the sequential traversal reaches them
    (kdop_tandem_traversal.h:40-151, 160-196): branch pairs row-major; per
    node pair the strict overlap test, then A's children outer and B's inner,
    or the one side that is not a leaf."""
    out = []

    def overlap(x, y):
        u, v = _aabb(x, box_a), _aabb(y, box_b)
        return not any(u[2 * k + 1] < v[2 * k] or v[2 * k + 1] < u[2 * k] for k in range(3))

    def rec(x, y):
        if not overlap(x, y):
            return
        xl, yl = isinstance(x, int), isinstance(y, int)
        if xl and yl:
            out.append((x, y))
        elif not xl and not yl:
            for a in x:
                for b in y:
                    rec(a, b)
        elif not xl:
            for a in x:
                rec(a, y)
        else:
            for b in y:
                rec(x, b)

    ra = [_aabb(t, box_a) for t in sa]
    rb = [_aabb(t, box_b) for t in sb]
    root_a = np.array([f(r[k] for r in ra) for k, f in enumerate([min, max] * 3)])
    root_b = np.array([f(r[k] for r in rb) for k, f in enumerate([min, max] * 3)])
    if any(root_a[2 * k + 1] < root_b[2 * k] or root_b[2 * k + 1] < root_a[2 * k] for k in range(3)):
        return out
    for x in sa:
        for y in sb:
            rec(x, y)
    return out


def _tet_boxes(mesh, X, Y, Z):
    P = np.stack([X, Y, Z], 1)[mesh.tets.astype(np.int64)]  # (nt, 4, 3), exact float32 coordinates
    lo, hi = P.min(1), P.max(1)
    return np.stack([lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1], lo[:, 2], hi[:, 2]], 1)


@pytest.mark.parametrize("case", ["all_overlap", "offset"])
def test_sequential_leaf_pair_order_is_the_reference_recursion(case):
    """The oracle's P1 list in sequential order (threads=1) equals an
    enumeration written from the reference's recursion alone: chunk trees
    from the median split, K = 6 volumes as exact per-tet AABBs and their
    unions (test_leaf_dop_k6_is_exact_aabb), strict overlap, branch pairs
    row-major.  A swapped loop nest, a reversed child order or a missing
    pruning test reorders or changes the list."""
    if case == "all_overlap":
        ma, mb = _star_mesh(2), _star_mesh(3)
        Xb, Yb, Zb = mb.X, mb.Y, mb.Z
    else:
        ma, mb = K.box_mesh(6, jitter=0.2, seed=5), K.box_mesh(5, jitter=0.2, seed=6)
        Xb, Yb, Zb = (mb.X + F32(0.55)).astype(F32), (mb.Y + F32(0.3)).astype(F32), mb.Z
    ta = O.Tree(6, 8000, ma.tets, ma.X, ma.Y, ma.Z)
    tb = O.Tree(6, 8000, mb.tets, mb.X, mb.Y, mb.Z)
    tb.refit(Xb, Yb, Zb)
    c, p1, _, _, _ = O.traverse(ta, _mesh_tuple(ma), tb, _mesh_tuple(mb, Xb, Yb, Zb), narrow=False, threads=1)
    La = _shape_closed_form(6, 8000, ma.nt)["L"]
    Lb = _shape_closed_form(6, 8000, mb.nt)["L"]
    want = _visit_order(_branch_trees(ma.nt, La), _branch_trees(mb.nt, Lb),
                        _tet_boxes(ma, ma.X, ma.Y, ma.Z), _tet_boxes(mb, Xb, Yb, Zb))
    assert len(want) > 500 and (case == "all_overlap" or len(want) < ma.nt * mb.nt // 4)
    assert c["leaf_pairs"] == len(want)
    assert [tuple(int(v) for v in r) for r in p1] == want


# --------------------------------------------------------------- surface map
def _surface_bits(S):
    return [(t, v) for t in range(S.shape[0]) for v in range(4) if (int(S[t]) >> v) & 1]


@pytest.mark.parametrize("n,jitter,reorder", [(1, 0.0, False), (2, 0.0, False), (3, 0.2, False), (4, 0.15, True)])
def test_surface_map_closed_form_on_box(n, jitter, reorder):
    """An n^3 box of 5-tet cubes has 6 n^2 boundary squares of 2 triangles:
    12 n^2 surface bits.  Bit v of tet t (mesh_array_compute_surface_map.h:
    41-76: the face opposite vertex v has no neighbour) must name a face
    whose three vertices lie on one boundary plane while vertex v does not;
    jitter moves interior vertices only, so the planes stay exact."""
    if reorder:
        m = K.box_mesh(n, jitter=jitter, seed=4)
        X, Y, Z, T, S = m.X, m.Y, m.Z, m.tets, m.surface
    else:
        X, Y, Z, T = K.box_mesh_raw(n, 1.0, jitter, 4)
        S = K.compute_surface_map(X.shape[0], T)
    bits = _surface_bits(S)
    assert len(bits) == 12 * n * n
    P = np.stack([X, Y, Z], 1)
    for t, v in bits:
        face = [int(T[t][q]) for q in range(4) if q != v]
        on = [ax for ax in range(3) for s in (-0.5, 0.5) if all(P[i][ax] == F32(s) for i in face)]
        assert len(on) == 1, (t, v, P[face])
        ax = on[0]
        assert P[int(T[t][v])][ax] != P[face[0]][ax]


def test_surface_map_two_tets_known_answer():
    """Two tets sharing the face {1, 2, 3}: in (0, 1, 2, 3) it lies opposite
    local vertex 0, in (4, 2, 1, 3) opposite local vertex 0 too, in
    (1, 2, 3, 4) opposite local vertex 3; every other face is on the surface."""
    T = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.uint32)
    assert list(K.compute_surface_map(5, T)) == [0b1110, 0b0111]
    T = np.array([[0, 1, 2, 3], [4, 2, 1, 3]], np.uint32)
    assert list(K.compute_surface_map(5, T)) == [0b1110, 0b1110]
    T = np.array([[0, 1, 2, 3], [5, 6, 7, 4]], np.uint32)  # no shared face
    assert list(K.compute_surface_map(8, T)) == [0b1111, 0b1111]


# ------------------------------------------------------------------- rotate
def test_rotate_matches_rotation_matrix():
    rng = np.random.default_rng(4)
    P = rng.normal(size=(500, 3)).astype(F32)
    for q in K.rng_quaternions(3, 0, 20):
        p = rng.normal(size=3).astype(F32)
        X, Y, Z = O.transform_f32(P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), p, q)
        qd = q.astype(np.float64)
        x, y, z, w = qd
        n2 = qd @ qd
        R = np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])
        want = P.astype(np.float64) @ R.T + p.astype(np.float64)
        got = np.stack([X, Y, Z], 1).astype(np.float64)
        scale = np.abs(P).max() * max(n2, 1.0) + np.abs(p).max()
        assert np.abs(got - want).max() <= 16 * np.finfo(F32).eps * scale
        Xd, Yd, Zd = O.transform_f64(P[:, 0], P[:, 1], P[:, 2], p, q)
        assert np.abs(np.stack([Xd, Yd, Zd], 1) - want).max() <= 1e-13 * scale


# ---------------------------------------------------------------------- SAT
def _lp_margin(A, B):
    """max s such that some x lies inside both tets with margin s (negative:
    separated).  Independent of the SAT code: scipy linprog over the 8 face
    half-spaces written with outward unit normals."""
    from scipy.optimize import linprog
    rows, rhs = [], []
    for T in (A, B):
        c = T.mean(0)
        for f in ((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)):
            p0, p1, p2 = T[list(f)]
            n = np.cross(p1 - p0, p2 - p0)
            n /= np.linalg.norm(n)
            if n @ (c - p0) > 0:
                n = -n
            rows.append(np.append(n, 1.0))
            rhs.append(n @ p0)
    res = linprog(c=[0, 0, 0, -1], A_ub=np.array(rows), b_ub=np.array(rhs), bounds=[(None, None)] * 3 + [(None, 10)])
    return -res.fun


def _random_tet(rng, center, size):
    while True:
        P = center + rng.normal(size=(4, 3)) * size
        v = np.dot(P[3] - P[0], np.cross(P[1] - P[0], P[2] - P[0]))
        if abs(v) > 0.05 * size ** 3:
            if v < 0:
                P[[2, 3]] = P[[3, 2]]
            return P


def test_sat_matches_lp_feasibility():
    rng = np.random.default_rng(7)
    decided = 0
    for _ in range(400):
        A = _random_tet(rng, np.zeros(3), 1.0)
        B = _random_tet(rng, rng.normal(size=3) * 1.2, 1.0)
        m = _lp_margin(A, B)
        got = O.overlap_tet_tet(A.astype(F32), B.astype(F32))
        if m > 1e-3:
            assert got, m
            decided += 1
        elif m < -1e-3:
            # separated with margin: the LP's dual gives no simple distance, so
            # use the true gap via the SAT in float64 on the same 44 axes
            assert not got or m > -1e-3
            decided += 1
    assert decided > 300


# ----------------------------------------------------------------- contacts
def test_contacts_known_answer_face_penetration():
    """B's top face pushes d into A's bottom face: normal (0,0,-1), depth -d.
    Candidates: B's three top vertices (inside A) and the three crossings of
    B's side edges with z = 0; projected onto the mid plane z = d/2 the
    crossing of the vertical edge coincides with its top vertex, leaving five
    contacts (contacts ...tetrahedron.h:543-681)."""
    d = 0.05
    A = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    B = np.array([[0.1, 0.1, d], [0.1, 0.3, d], [0.3, 0.1, d], [0.1, 0.1, -1.0]], np.float64)
    sa, sb = 1 << 3, 1 << 3  # faces opposite vertex 3 are on the surface
    for dt in (np.float32, np.float64):
        n, depth, face, pts, ncand = O.contacts_pair(A, B, sa, sb, dtype=dt)
        tol = 1e-6 if dt == np.float32 else 1e-12
        assert face == 3
        np.testing.assert_allclose(n, [0, 0, -1], atol=tol)
        assert abs(depth + d) <= tol
        assert pts.shape[0] == 5 and ncand == 6
        apex = B[3]
        t = d / (1 + d)
        cross = [B[i] + (apex - B[i]) * t for i in (1, 2)]
        want = np.array([[0.1, 0.1, d / 2], [0.1, 0.3, d / 2], [0.3, 0.1, d / 2],
                         [cross[0][0], cross[0][1], d / 2], [cross[1][0], cross[1][1], d / 2]])
        got = pts[np.lexsort((pts[:, 1], pts[:, 0]))]
        np.testing.assert_allclose(got, want[np.lexsort((want[:, 1], want[:, 0]))], atol=4 * tol)


def test_contacts_rejected_without_opposing_surfaces():
    A = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    B = np.array([[0.1, 0.1, 0.05], [0.1, 0.3, 0.05], [0.3, 0.1, 0.05], [0.1, 0.1, -1.0]])
    # surface faces that face the same way cannot be most opposing
    n, depth, face, pts, ncand = O.contacts_pair(A, B, 1 << 3, 1 << 0, dtype=np.float64)
    assert face == -1 and pts.shape[0] == 0


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_float32_contacts_agree_with_frozen_float64(jitter):
    """Every float32 contact vs the FP64 restatement of the same pair with the
    float run's discrete choices frozen: points, normal and depth within 1e-5
    (relative to max(|x|, 1)).  Unfrozen FP64 may legitimately differ on the
    unjittered grid, where vertices lie exactly on face planes."""
    from kdop_amd import workloads as W
    wl = W.single_pair(n=10, K=8) if jitter == 0.0 else W.large_pair(n=10, K=8, seed=21, jitter=jitter)
    recs, PA, PB, ma, mb = narrow_records(wl)
    assert len(recs) > 5
    worst = 0.0
    for ta_i, tb_i, r in recs:
        n64, d64, p64 = O.frozen_pair_f64(PA[ma.tets[ta_i]], PB[mb.tets[tb_i]], r["face"], r["codes"], r["uniq"],
                                          r["imin"], r["imax"])
        assert p64.shape == r["pts"].shape
        worst = max(worst, np.abs(r["n"] - n64).max(), abs(r["depth"] - d64), np.abs(r["pts"] - p64).max())
    assert worst < 1e-5, worst


def test_frozen_gradient_matches_central_differences():
    """Analytic FP64 gradient of Phi (frozen choices) vs central differences,
    h = 1e-3 on every pose component of both bodies (the quaternion is used
    literally, not renormalised)."""
    from kdop_amd import workloads as W
    wl = W.large_pair(n=8, K=14, seed=17, jitter=0.1)
    recs, PA, PB, ma, mb = narrow_records(wl)
    assert len(recs) > 3
    R = np.zeros(len(recs), dtype=[("ta", np.uint32), ("tb", np.uint32), ("face", np.uint8), ("nuniq", np.uint8),
                                   ("cmin", np.uint8), ("cmax", np.uint8)])
    for i, (ta_i, tb_i, r) in enumerate(recs):
        R[i] = (ta_i, tb_i, r["face"], r["pts"].shape[0], r["codes"][r["imin"]], r["codes"][r["imax"]])
    pa, qa = wl.poses[0].p.astype(np.float64), wl.poses[0].q.astype(np.float64)
    pb, qb = wl.poses[1].p.astype(np.float64), wl.poses[1].q.astype(np.float64)
    phi, gpa, gqa, gpb, gqb = O.phi_frozen_f64(R, ma, (pa, qa), mb, (pb, qb), grad=True)
    want_phi = sum(r["pts"].shape[0] * r["depth"] for _, _, r in recs)
    assert abs(phi - want_phi) <= 1e-5 * max(1.0, abs(want_phi))
    grad = np.concatenate([gpa, gqa, gpb, gqb])
    h = 1e-3
    base = [pa, qa, pb, qb]
    fd = []
    for blk in range(4):
        for j in range(base[blk].shape[0]):
            vals = []
            for sgn in (1, -1):
                x = [v.copy() for v in base]
                x[blk][j] += sgn * h
                vals.append(O.phi_frozen_f64(R, ma, (x[0], x[1]), mb, (x[2], x[3])))
            fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    scale = max(np.abs(fd).max(), 1e-6)
    assert np.abs(grad - fd).max() <= 1e-4 * scale, (grad, fd)
