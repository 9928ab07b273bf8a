"""Pins of the oracle's raycast (kdop_raycast.h:102-155 with
raycast/geometry_raycast_dop.h and geometry_raycast_tetrahedron.h),
independent of its code: closed forms on one tetrahedron, the traversal's
result against the best single-tetrahedron raycast over every tet (brute
force, so the tree's pruning loses no hit), and every hit distance against a
float64 ray-triangle intersection.  CPU only."""
import numpy as np
import pytest

import kdop_amd as K
from oracle import oracle as O

F32 = np.float32
FMAX = np.finfo(F32).max


def _tree(mesh, k=8, chunk_bytes=8000):
    t = O.Tree(k, chunk_bytes, mesh.tets, mesh.X, mesh.Y, mesh.Z)
    t.refit(mesh.X, mesh.Y, mesh.Z)
    return t, (mesh.tets, mesh.X, mesh.Y, mesh.Z, mesh.surface)


def _unit_tet():
    X, Y, Z = F32([0, 1, 0, 0]), F32([0, 0, 1, 0]), F32([0, 0, 0, 1])
    T = np.array([[0, 1, 2, 3]], np.uint32)
    return K.Mesh(X, Y, Z, T, K.compute_surface_map(4, T))


def test_single_tet_closed_forms():
    t, m = _tree(_unit_tet())
    o = F32([[0.2, 0.2, 5.0], [0.2, 0.2, 5.0], [0.25, 0.25, 0.25], [5.0, 0.1, 0.1]])
    d = F32([[0, 0, -1], [0, 0, 1], [1, 1, 1], [-2, 0, 0]])
    hit, length, flag = O.raycast(t, m, o, d)
    # straight down onto x + y + z = 1 at (0.2, 0.2, 0.6)
    assert flag[0] and abs(length[0] - 4.4) < 1e-6 and np.allclose(hit[0], [0.2, 0.2, 0.6], atol=1e-6)
    # pointing away: a miss reports (0, 0, 0) and FLT_MAX
    assert not flag[1] and length[1] == FMAX and not hit[1].any()
    # an origin inside the root DOP never hits (the root test needs an entry in front of the origin)
    assert not flag[2] and length[2] == FMAX
    # along -x (a direction of length 2, normalised) onto the same face at x = 0.8
    assert flag[3] and abs(length[3] - 4.2) < 1e-6 and np.allclose(hit[3], [0.8, 0.1, 0.1], atol=1e-6)


def _rays(rng, n, center, radius):
    """Origins on a sphere around the body, aimed at points inside it."""
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = center + radius * u
    target = center + rng.uniform(-0.45, 0.45, size=(n, 3))
    return o.astype(F32), (target - o).astype(F32)


def _tri_distance64(o, d, a, b, c):
    """float64 ray-triangle intersection (Moller-Trumbore), either side; None on a miss."""
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-14:
        return None
    s = o - a
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    if u < -1e-9 or v < -1e-9 or u + v > 1 + 1e-9 or t < 0:
        return None
    return t


def test_traversal_equals_brute_force_over_tets():
    rng = np.random.default_rng(3)
    mesh = K.box_mesh(3, size=1.0, jitter=0.1, seed=5)
    t, m = _tree(mesh)
    o, d = _rays(rng, 60, np.zeros(3), 2.5)
    hit, length, flag = O.raycast(t, m, o, d)
    # every tet on its own (a one-tet tree, surface bits of the whole mesh)
    best = np.full(o.shape[0], FMAX, F32)
    for ti in range(mesh.tets.shape[0]):
        T = mesh.tets[ti:ti + 1].copy()
        sub = K.Mesh(mesh.X, mesh.Y, mesh.Z, T, mesh.surface[ti:ti + 1].copy())
        ts = O.Tree(8, 8000, sub.tets, sub.X, sub.Y, sub.Z)
        ts.refit(sub.X, sub.Y, sub.Z)
        _, l1, f1 = O.raycast(ts, (sub.tets, sub.X, sub.Y, sub.Z, sub.surface), o, d)
        best = np.where(f1 & (l1 < best), l1, best)
    np.testing.assert_array_equal(length, best)
    assert flag.sum() > 30  # most rays hit
    # hit distances against float64 geometry (surface faces only, the nearest)
    V = np.stack([mesh.X, mesh.Y, mesh.Z], 1).astype(np.float64)
    faces = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
    for i in np.nonzero(flag)[0]:
        oo = o[i].astype(np.float64)
        dd = d[i].astype(np.float64)
        dd /= np.linalg.norm(dd)
        ref = np.inf
        for ti, tet in enumerate(mesh.tets):
            for f, (a, b, c) in enumerate(faces):
                if not (mesh.surface[ti] >> f) & 1:
                    continue
                tt = _tri_distance64(oo, dd, V[tet[a]], V[tet[b]], V[tet[c]])
                if tt is not None:
                    ref = min(ref, tt)
        assert abs(length[i] - ref) <= 1e-5 * max(1.0, ref), (i, length[i], ref)
