"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker of the HIP library; nothing in the product imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64

CONTACT_DTYPE = np.dtype([("ta", np.uint32), ("tb", np.uint32), ("p", np.float32, (3,)), ("n", np.float32, (3,)),
                          ("d", np.float32)])

COUNTER_NAMES = ("tests", "branch_tests", "descent_tests", "branch_overlaps", "leaf_pairs", "surface_pairs",
                 "contacts", "narrow_pairs", "root_tests")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_dir_table_f32.argtypes = [ctypes.c_int, _p]
        L.orc_tree_shape.argtypes = [ctypes.c_int, _u64, _u64, _p]
        L.orc_mesh_reorder.argtypes = [_u32, _p, _p, _p, _u32, _p, _p, _p, _p, _p]
        L.orc_tree_build.restype = _p
        L.orc_tree_build.argtypes = [ctypes.c_int, _u64, _u32, _p, _p, _p, _p]
        L.orc_tree_free.argtypes = [_p]
        L.orc_tree_refit.argtypes = [_p, _p, _p, _p, _p]
        L.orc_tree_levels.argtypes = [_p, _p]
        L.orc_tree_root.argtypes = [_p, _p]
        L.orc_tree_count_nodes.restype = _u64
        L.orc_tree_count_nodes.argtypes = [_p]
        L.orc_tree_export.restype = _u64
        L.orc_tree_export.argtypes = [_p, _p, _p, _p]
        L.orc_transform_f32.argtypes = [_u32, _p, _p, _p, _p, _p, _p, _p, _p]
        L.orc_transform_f64.argtypes = [_u32, _p, _p, _p, _p, _p, _p, _p, _p]
        L.orc_make_dop_f32.argtypes = [ctypes.c_int, _p, _p]
        L.orc_traverse.argtypes = [_p, ctypes.c_int, _p, _p, _p, _p, _p, _p, ctypes.c_int, _p, _p, _p, _p, _p,
                                   ctypes.c_int, ctypes.c_int, _p, _u64, _p, _u64, _p, _u64, _p, _p]
        L.orc_traverse_rows.argtypes = [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _u64, _u64, ctypes.c_int, _p, _p]
        L.orc_traverse_rows_mt.argtypes = [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _u64, _u64, ctypes.c_int,
                                           ctypes.c_int, _p, _p]
        L.orc_raycast.argtypes = [_p, _p, _p, _p, _p, _p, _u32, _p, _p, _p, _p, _p]
        L.orc_single_traversal.restype = _u64
        L.orc_single_traversal.argtypes = [_p, _p, _p, _p, _p, _p, _u32, _p, ctypes.c_int, _p, _u64]
        L.orc_prox_reduce.restype = _u64
        L.orc_prox_reduce.argtypes = [_u64, _p, _p, _p, _p, _p, _p]
        L.orc_brute_pairs.restype = _u64
        L.orc_brute_pairs.argtypes = [_p, _p, _p, _u64]
        L.orc_contacts_pair_f32.argtypes = [_p, _p, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, ctypes.c_int, _p]
        L.orc_contacts_pair_f64.argtypes = [_p, _p, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, ctypes.c_int, _p]
        L.orc_overlap_tet_tet_f32.argtypes = [_p, _p]
        L.orc_contacts_pair_codes_f32.argtypes = [_p, _p, ctypes.c_int, ctypes.c_int, _p, _p, _p, _p, ctypes.c_int, _p,
                                                  _p, _p, _p]
        L.orc_frozen_pair_f64.argtypes = [_p, _p, ctypes.c_int, ctypes.c_int, _p, _p, ctypes.c_int, ctypes.c_int, _p, _p, _p]
        L.orc_phi_frozen_f64.restype = ctypes.c_double
        L.orc_phi_frozen_f64.argtypes = [_u32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                         _p]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def dir_table(K):
    out = np.zeros((13, 3), np.float32)
    n = lib().orc_dir_table_f32(K, _ptr(out))
    return out[:n]


def mesh_reorder(X, Y, Z, tets):
    """Morton reorder (kdop_mesh_reorder.h:172-353, stable on equal codes): (X, Y, Z, tets) reordered."""
    X, Y, Z = (np.ascontiguousarray(a, np.float32) for a in (X, Y, Z))
    T = np.ascontiguousarray(np.asarray(tets, np.uint32).reshape(-1, 4))
    nv, nt = X.shape[0], T.shape[0]
    Xo, Yo, Zo = (np.empty(nv, np.float32) for _ in range(3))
    To = np.empty((nt, 4), np.uint32)
    if lib().orc_mesh_reorder(nv, _ptr(X), _ptr(Y), _ptr(Z), nt, _ptr(T), _ptr(Xo), _ptr(Yo), _ptr(Zo), _ptr(To)):
        raise ValueError("orc_mesh_reorder: empty mesh")
    return Xo, Yo, Zo, To


def tree_shape(K, chunk_bytes, nt):
    out = np.zeros(5, np.uint64)
    if lib().orc_tree_shape(K, chunk_bytes, nt, _ptr(out)):
        return None
    return dict(zip(("node_bytes", "n_perfect", "L", "C", "H"), (int(v) for v in out)))


def transform_f32(X0, Y0, Z0, p, q):
    n = X0.shape[0]
    X, Y, Z = (np.empty(n, np.float32) for _ in range(3))
    p = np.ascontiguousarray(p, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    lib().orc_transform_f32(n, _ptr(X0), _ptr(Y0), _ptr(Z0), _ptr(p), _ptr(q), _ptr(X), _ptr(Y), _ptr(Z))
    return X, Y, Z


def transform_f64(X0, Y0, Z0, p, q):
    n = X0.shape[0]
    X0, Y0, Z0 = (np.ascontiguousarray(a, np.float64) for a in (X0, Y0, Z0))
    X, Y, Z = (np.empty(n, np.float64) for _ in range(3))
    p = np.ascontiguousarray(p, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    lib().orc_transform_f64(n, _ptr(X0), _ptr(Y0), _ptr(Z0), _ptr(p), _ptr(q), _ptr(X), _ptr(Y), _ptr(Z))
    return X, Y, Z


def make_dop(K, pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(4, 3)
    out = np.empty(K, np.float32)
    lib().orc_make_dop_f32(K, _ptr(pts), _ptr(out))
    return out


class Tree:
    """Reference-layout chunked kDOP tree (kdop::make_tree, sequential)."""

    def __init__(self, K, chunk_bytes, tets, X, Y, Z):
        self.K = K
        self.tets = np.ascontiguousarray(tets, np.uint32)
        X, Y, Z = (np.ascontiguousarray(a, np.float32) for a in (X, Y, Z))
        self.h = lib().orc_tree_build(K, chunk_bytes, self.tets.shape[0], _ptr(self.tets), _ptr(X), _ptr(Y), _ptr(Z))
        if not self.h:
            raise ValueError("oracle make_tree rejected the input")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def refit(self, X, Y, Z):
        X, Y, Z = (np.ascontiguousarray(a, np.float32) for a in (X, Y, Z))
        lib().orc_tree_refit(self.h, _ptr(self.tets), _ptr(X), _ptr(Y), _ptr(Z))

    def levels(self):
        out = np.zeros(64, np.uint64)
        n = lib().orc_tree_levels(self.h, _ptr(out))
        return [int(v) for v in out[:n]]

    def root(self):
        out = np.empty(self.K, np.float32)
        lib().orc_tree_root(self.h, _ptr(out))
        return out

    def export(self):
        n = lib().orc_tree_count_nodes(self.h)
        info = np.empty((n, 6), np.uint64)
        rng = np.empty((n, 2), np.uint64)
        vol = np.empty((n, self.K), np.float32)
        lib().orc_tree_export(self.h, _ptr(info), _ptr(rng), _ptr(vol))
        return dict(info=info, range=rng, vol=vol)


def traverse(tree_a, mesh_a, tree_b, mesh_b, narrow=True, threads=1, want_p1=True, want_p2=True, want_contacts=True,
             cap=None):
    """Sequential CPU tandem traversal (kdop_tandem_traversal.h:160-226).

    mesh_*: (tets, X, Y, Z, surface).  Returns (counters dict, p1, p2, contacts, seconds).
    """
    L = lib()
    ta, Xa, Ya, Za, Sa = (np.ascontiguousarray(a) for a in mesh_a)
    tb, Xb, Yb, Zb, Sb = (np.ascontiguousarray(a) for a in mesh_b)
    cnt = np.zeros(9, np.uint64)
    sec = ctypes.c_double()
    # counting pass to size the outputs (cheap relative to a second pass only when small)
    if cap is None:
        L.orc_traverse(tree_a.h, 0, _ptr(ta), _ptr(Xa), _ptr(Ya), _ptr(Za), _ptr(Sa),
                       tree_b.h, 0, _ptr(tb), _ptr(Xb), _ptr(Yb), _ptr(Zb), _ptr(Sb),
                       int(narrow), threads, None, 0, None, 0, None, 0, _ptr(cnt), ctypes.byref(sec))
        c1, c2, cc = int(cnt[4]), int(cnt[5]), int(cnt[6])
    else:
        c1 = c2 = cc = cap
    p1 = np.empty((max(c1, 1), 2), np.uint32) if want_p1 else None
    p2 = np.empty((max(c2, 1), 2), np.uint32) if want_p2 else None
    con = np.empty(max(cc, 1), CONTACT_DTYPE) if (want_contacts and narrow) else None
    L.orc_traverse(tree_a.h, 0, _ptr(ta), _ptr(Xa), _ptr(Ya), _ptr(Za), _ptr(Sa),
                   tree_b.h, 0, _ptr(tb), _ptr(Xb), _ptr(Yb), _ptr(Zb), _ptr(Sb),
                   int(narrow), threads, _ptr(p1), c1 if want_p1 else 0, _ptr(p2), c2 if want_p2 else 0,
                   _ptr(con), cc if con is not None else 0, _ptr(cnt), ctypes.byref(sec))
    counters = dict(zip(COUNTER_NAMES, (int(v) for v in cnt)))
    n1, n2, nc = counters["leaf_pairs"], counters["surface_pairs"], counters["contacts"]
    return (counters, p1[:n1] if p1 is not None else None, p2[:n2] if p2 is not None else None,
            con[:nc] if con is not None else None, sec.value)


def traverse_rows(tree_a, mesh_a, tree_b, mesh_b, row0, row1, narrow=True):
    """Single-threaded timing sample: branch rows [row0, row1) of A against all of B.
    Returns (counters dict, seconds)."""
    ta, Xa, Ya, Za, Sa = (np.ascontiguousarray(a) for a in mesh_a)
    tb, Xb, Yb, Zb, Sb = (np.ascontiguousarray(a) for a in mesh_b)
    cnt = np.zeros(9, np.uint64)
    sec = ctypes.c_double()
    lib().orc_traverse_rows(tree_a.h, _ptr(ta), _ptr(Xa), _ptr(Ya), _ptr(Za), _ptr(Sa),
                            tree_b.h, _ptr(tb), _ptr(Xb), _ptr(Yb), _ptr(Zb), _ptr(Sb),
                            int(row0), int(row1), int(narrow), _ptr(cnt), ctypes.byref(sec))
    return dict(zip(COUNTER_NAMES, (int(v) for v in cnt))), sec.value


def traverse_rows_mt(tree_a, mesh_a, tree_b, mesh_b, row0, row1, threads, narrow=True):
    """All-cores timing sample: branch rows [row0, row1) of A against all of B on
    `threads` threads taking rows dynamically.  Returns (counters dict, seconds)."""
    ta, Xa, Ya, Za, Sa = (np.ascontiguousarray(a) for a in mesh_a)
    tb, Xb, Yb, Zb, Sb = (np.ascontiguousarray(a) for a in mesh_b)
    cnt = np.zeros(9, np.uint64)
    sec = ctypes.c_double()
    lib().orc_traverse_rows_mt(tree_a.h, _ptr(ta), _ptr(Xa), _ptr(Ya), _ptr(Za), _ptr(Sa),
                               tree_b.h, _ptr(tb), _ptr(Xb), _ptr(Yb), _ptr(Zb), _ptr(Sb),
                               int(row0), int(row1), int(narrow), int(threads), _ptr(cnt), ctypes.byref(sec))
    return dict(zip(COUNTER_NAMES, (int(v) for v in cnt))), sec.value


def raycast(tree, mesh, origins, directions):
    """kdop::raycast (kdop_raycast.h:102-155) of n rays against a refitted tree:
    (hit points (n, 3), lengths (n,), hit flags (n,) bool)."""
    t, X, Y, Z, S = (np.ascontiguousarray(a) for a in mesh)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
    n = o.shape[0]
    hits = np.zeros((n, 3), np.float32)
    lengths = np.zeros(n, np.float32)
    flags = np.zeros(n, np.uint8)
    lib().orc_raycast(tree.h, _ptr(t), _ptr(X), _ptr(Y), _ptr(Z), _ptr(S), n, _ptr(o), _ptr(d), _ptr(hits), _ptr(lengths), _ptr(flags))
    return hits, lengths, flags.astype(bool)


SPHERE_CONTACT_DTYPE = np.dtype([("sphere", np.uint32), ("tet", np.uint32), ("p", np.float32, (3,)), ("n", np.float32, (3,)),
                                 ("depth", np.float32)])


def single_traversal(tree, mesh, spheres, flip=False):
    """kdop::single_traversal (kdop_single_traversal.h:123-161) of spheres (n, 4: x, y, z, radius)
    against a refitted tree: contact records grouped by sphere, each sphere's in report order."""
    t, X, Y, Z, S = (np.ascontiguousarray(a) for a in mesh)
    sp = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
    n = int(lib().orc_single_traversal(tree.h, _ptr(t), _ptr(X), _ptr(Y), _ptr(Z), _ptr(S), sp.shape[0], _ptr(sp), int(flip), None, 0))
    out = np.zeros(n, SPHERE_CONTACT_DTYPE)
    if n:
        lib().orc_single_traversal(tree.h, _ptr(t), _ptr(X), _ptr(Y), _ptr(Z), _ptr(S), sp.shape[0], _ptr(sp), int(flip), _ptr(out), n)
    return out


def prox_reduce(body_i, body_j, p, n, depth):
    """PROX contact reduction (prox_collision_detection.h:340-345: std::sort, std::unique
    over prox::ContactPoint): the input positions of the kept contacts, in sorted order."""
    bi = np.ascontiguousarray(body_i, np.uint32)
    bj = np.ascontiguousarray(body_j, np.uint32)
    P = np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1, 3))
    N = np.ascontiguousarray(np.asarray(n, np.float32).reshape(-1, 3))
    D = np.ascontiguousarray(depth, np.float32)
    m = bi.shape[0]
    assert bj.shape[0] == m and P.shape[0] == m and N.shape[0] == m and D.shape[0] == m
    out = np.zeros(max(m, 1), np.uint32)
    k = int(lib().orc_prox_reduce(m, _ptr(bi), _ptr(bj), _ptr(P), _ptr(N), _ptr(D), _ptr(out)))
    return out[:k]


def brute_pairs(tree_a, tree_b, cap):
    out = np.empty((max(cap, 1), 2), np.uint32)
    n = lib().orc_brute_pairs(tree_a.h, tree_b.h, _ptr(out), cap)
    return out[:min(n, cap)], int(n)


def contacts_pair(pa, pb, sa, sb, dtype=np.float32):
    """Narrow phase of one tet pair: (n, depth, face, points (k x 3), ncand)."""
    pa = np.ascontiguousarray(pa, dtype).reshape(12)
    pb = np.ascontiguousarray(pb, dtype).reshape(12)
    n = np.zeros(3, dtype)
    d = np.zeros(1, dtype)
    face = ctypes.c_int()
    ncand = ctypes.c_int()
    pts = np.zeros((56, 3), dtype)
    f = lib().orc_contacts_pair_f32 if dtype == np.float32 else lib().orc_contacts_pair_f64
    k = f(_ptr(pa), _ptr(pb), int(sa), int(sb), _ptr(n), _ptr(d), ctypes.byref(face), _ptr(pts), 56, ctypes.byref(ncand))
    return n, float(d[0]), face.value, pts[:k], ncand.value


def overlap_tet_tet(pa, pb):
    pa = np.ascontiguousarray(pa, np.float32).reshape(12)
    pb = np.ascontiguousarray(pb, np.float32).reshape(12)
    return bool(lib().orc_overlap_tet_tet_f32(_ptr(pa), _ptr(pb)))


def contacts_pair_codes(pa, pb, sa, sb):
    """Float32 narrow phase of one pair plus its discrete choices:
    dict(n, depth, face, pts, codes, uniq, imin, imax)."""
    pa = np.ascontiguousarray(pa, np.float32).reshape(12)
    pb = np.ascontiguousarray(pb, np.float32).reshape(12)
    n = np.zeros(3, np.float32)
    d = np.zeros(1, np.float32)
    face = ctypes.c_int()
    ncand = ctypes.c_int()
    pts = np.zeros((56, 3), np.float32)
    codes = np.zeros(56, np.uint8)
    uniq = np.zeros(56, np.uint8)
    mm = np.zeros(2, np.int32)
    k = lib().orc_contacts_pair_codes_f32(_ptr(pa), _ptr(pb), int(sa), int(sb), _ptr(n), _ptr(d), ctypes.byref(face),
                                          _ptr(pts), 56, ctypes.byref(ncand), _ptr(codes), _ptr(uniq), _ptr(mm))
    nc = ncand.value
    return dict(n=n, depth=float(d[0]), face=face.value, pts=pts[:k], codes=codes[:nc], uniq=uniq[:nc],
                imin=int(mm[0]), imax=int(mm[1]))


def frozen_pair_f64(pa, pb, face, codes, uniq, imin, imax):
    """FP64 evaluation of one pair with the float run's discrete choices."""
    pa = np.ascontiguousarray(pa, np.float64).reshape(12)
    pb = np.ascontiguousarray(pb, np.float64).reshape(12)
    codes = np.ascontiguousarray(codes, np.uint8)
    uniq = np.ascontiguousarray(uniq, np.uint8)
    n = np.zeros(3)
    d = np.zeros(1)
    pts = np.zeros((56, 3))
    k = lib().orc_frozen_pair_f64(_ptr(pa), _ptr(pb), int(face), codes.shape[0], _ptr(codes), _ptr(uniq), int(imin),
                                  int(imax), _ptr(n), _ptr(d), _ptr(pts))
    return n, float(d[0]), pts[:k]


def phi_frozen_f64(recs, mesh_a, pose_a, mesh_b, pose_b, grad=False):
    """Phi = sum nuniq * depth over narrow records (fields ta, tb, face, nuniq,
    cmin, cmax) of one body pair, in FP64 from material coordinates and poses
    (p[3], q[4] as x y z w).  With grad: (phi, gpA, gqA, gpB, gqB)."""
    cols = [np.ascontiguousarray(recs[f], dt) for f, dt in
            (("ta", np.uint32), ("tb", np.uint32), ("face", np.uint8), ("nuniq", np.uint8), ("cmin", np.uint8),
             ("cmax", np.uint8))]
    def m(mesh):
        return (np.ascontiguousarray(mesh.tets, np.uint32),) + tuple(np.ascontiguousarray(a, np.float64)
                                                                    for a in (mesh.X, mesh.Y, mesh.Z))
    ta, Xa, Ya, Za = m(mesh_a)
    tb, Xb, Yb, Zb = m(mesh_b)
    pa, qa, pb, qb = (np.ascontiguousarray(v, np.float64) for v in (pose_a[0], pose_a[1], pose_b[0], pose_b[1]))
    g = np.zeros(14) if grad else None
    phi = lib().orc_phi_frozen_f64(cols[0].shape[0], *(_ptr(c) for c in cols), _ptr(ta), _ptr(Xa), _ptr(Ya), _ptr(Za),
                                   _ptr(pa), _ptr(qa), _ptr(tb), _ptr(Xb), _ptr(Yb), _ptr(Zb), _ptr(pb), _ptr(qb), _ptr(g))
    if not grad:
        return phi
    return phi, g[0:3], g[3:7], g[7:10], g[10:14]
