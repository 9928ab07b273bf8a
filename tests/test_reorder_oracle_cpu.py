"""Pins of the oracle's Morton mesh reorder (kdop_mesh_reorder.h:172-353),
independent of its own code: a closed form on point clusters computed with
Python integers, the invariants every reorder must keep (a permutation of
the tets, the same mesh, codes in order with ties in input order, vertices
numbered in first-touch order), and agreement with the library's host
reorder.  CPU only."""
import numpy as np
import pytest

import kdop_amd as K
from oracle import oracle as O

F32 = np.float32


def _interleave(x, y, z, bits=21):
    c = 0
    for k in range(bits):
        c |= ((x >> k) & 1) << (3 * k)
        c |= ((y >> k) & 1) << (3 * k + 1)
        c |= ((z >> k) & 1) << (3 * k + 2)
    return c


def _cluster_mesh(points):
    """One degenerate tet per point (its four vertices all at the point): the
    centroid is the point itself, exactly."""
    P = np.asarray(points, F32)
    n = P.shape[0]
    X = np.repeat(P[:, 0], 4).astype(F32)
    Y = np.repeat(P[:, 1], 4).astype(F32)
    Z = np.repeat(P[:, 2], 4).astype(F32)
    T = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    return X, Y, Z, T


def _expected_order(points, lo, scale):
    codes = []
    for i, (x, y, z) in enumerate(points):
        q = [int(F32(F32(c) - F32(l)) * F32(scale)) for c, l in zip((x, y, z), lo)]
        codes.append((_interleave(*q), i))
    return [i for _, i in sorted(codes)]


@pytest.mark.parametrize("shift", [0.0, -10.0])
def test_grid_points_closed_form(shift):
    """Integer grid points in shuffled order; with shift -10 every centroid
    is negative and the maximum stays at FLT_MIN, as the reference's
    numeric_limits<float>::min() start makes it."""
    rng = np.random.default_rng(5)
    pts = [(i + shift, j + shift, k + shift) for i in range(4) for j in range(4) for k in range(4)]
    pts = [pts[i] for i in rng.permutation(len(pts))]
    lo = [min(p[a] for p in pts) for a in range(3)]
    hi = [max(max(p[a] for p in pts), float(np.finfo(F32).tiny)) for a in range(3)]
    scale = min(int(F32(2097151) / F32(F32(h) - F32(l))) for l, h in zip(lo, hi))
    X, Y, Z, T = _cluster_mesh(pts)
    Xo, Yo, Zo, To = O.mesh_reorder(X, Y, Z, T)
    order = _expected_order(pts, lo, scale)
    np.testing.assert_array_equal(Xo[To[:, 0]], X[T[order, 0]])
    np.testing.assert_array_equal(Yo[To[:, 0]], Y[T[order, 0]])
    np.testing.assert_array_equal(Zo[To[:, 0]], Z[T[order, 0]])
    # vertices in first-touch order: tet i names 4i .. 4i+3
    np.testing.assert_array_equal(To.reshape(-1), np.arange(4 * len(pts), dtype=np.uint32))


def _check_invariants(X, Y, Z, T, out):
    Xo, Yo, Zo, To = out
    nv, nt = X.shape[0], T.shape[0]
    # first-touch numbering: the first appearances in To are 0, 1, 2, ...
    flat = To.reshape(-1)
    _, first = np.unique(flat, return_index=True)
    np.testing.assert_array_equal(flat[np.sort(first)], np.arange(len(first), dtype=np.uint32))
    used = len(first)
    # the same mesh: every output tet is an input tet (as coordinate tuples), one to one
    key = lambda x, y, z, t: sorted(tuple(np.stack([x[t], y[t], z[t]], 1).reshape(-1).view(np.uint32)) for t in t)
    assert key(Xo, Yo, Zo, To) == key(X, Y, Z, T)
    # vertices no tet names follow in input order
    unused = np.setdiff1d(np.arange(nv), T.reshape(-1))
    np.testing.assert_array_equal(Xo[used:], X[unused])
    assert used + len(unused) == nv and nt == To.shape[0]


def test_invariants_and_host_agreement():
    rng = np.random.default_rng(11)
    X, Y, Z, T = K.box_mesh_raw(5, size=1.0, jitter=0.1, seed=2)
    T = T[rng.permutation(T.shape[0])]
    # two unused vertices at the end
    X = np.concatenate([X, F32([3.0, -3.0])]); Y = np.concatenate([Y, F32([0.5, 0.5])]); Z = np.concatenate([Z, F32([0.0, 1.0])])
    out = O.mesh_reorder(X, Y, Z, T)
    _check_invariants(X, Y, Z, T, out)
    host = K.mesh_reorder(X, Y, Z, T)
    for a, b in zip(out, host):
        assert a.tobytes() == b.tobytes()


def test_equal_codes_keep_input_order():
    """Different tets with one common centroid (the origin) keep their input
    order; a tet further out comes after them."""
    V = np.array([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
                  (2, 2, 2), (3, 2, 2), (2, 3, 2), (2, 2, 3)], F32)
    T = np.array([[4, 5, 0, 1],     # centroid 0, first
                  [6, 7, 8, 9],     # centroid (2.25, 2.25, 2.25)
                  [0, 1, 2, 3],     # centroid 0, second
                  [2, 3, 4, 5]],    # centroid 0, third
                 np.uint32)
    Xo, Yo, Zo, To = O.mesh_reorder(V[:, 0], V[:, 1], V[:, 2], T)
    got = [tuple(map(tuple, np.stack([Xo[t], Yo[t], Zo[t]], 1))) for t in To]
    want = [tuple(map(tuple, V[t])) for t in T[[0, 2, 3, 1]]]
    assert got == want
