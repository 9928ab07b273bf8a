"""Pins of the oracle's single traversal (kdop_single_traversal.h:123-161
with contacts/geometry_contacts_sphere_tetrahedron.h), independent of its
code: closed forms for each Voronoi region of one tetrahedron, and the
traversal's contacts against every tet on its own (brute force, so the
tree's pruning loses none).  CPU only."""
import numpy as np
import pytest

import kdop_amd as K
from oracle import oracle as O

F32 = np.float32
S3 = np.sqrt(3.0)


def _unit_tet():
    X, Y, Z = F32([0, 1, 0, 0]), F32([0, 0, 1, 0]), F32([0, 0, 0, 1])
    T = np.array([[0, 1, 2, 3]], np.uint32)
    return K.Mesh(X, Y, Z, T, K.compute_surface_map(4, T))


def _tree(mesh, k=8):
    t = O.Tree(k, 8000, mesh.tets, mesh.X, mesh.Y, mesh.Z)
    t.refit(mesh.X, mesh.Y, mesh.Z)
    return t, (mesh.tets, mesh.X, mesh.Y, mesh.Z, mesh.surface)


# (center, radius, expected point, expected normal (flip = False), expected depth), float64
CASES = [
    ((-0.3, -0.3, -0.3), 0.6, (0, 0, 0), (1 / S3, 1 / S3, 1 / S3), np.sqrt(0.27) - 0.6),   # vertex 0
    ((0.5, -0.3, -0.3), 0.5, (0.5, 0, 0), (0, 1 / np.sqrt(2), 1 / np.sqrt(2)), np.sqrt(0.18) - 0.5),  # edge 0-1
    ((0.2, 0.2, -0.3), 0.5, (0.2, 0.2, 0), (0, 0, 1), -0.2),                                 # face z = 0
    ((0.1, 0.1, 0.2), 0.05, (0, 0.1, 0.2), (1, 0, 0), -0.1),                                 # inside, nearest face x = 0
]


@pytest.mark.parametrize("case", range(len(CASES)), ids=["vertex", "edge", "face", "inside"])
@pytest.mark.parametrize("flip", [False, True])
def test_voronoi_regions_closed_form(case, flip):
    c, r, p, n, d = CASES[case]
    t, m = _tree(_unit_tet())
    out = O.single_traversal(t, m, [[*c, r]], flip=flip)
    assert out.shape[0] == 1 and out[0]["tet"] == 0 and out[0]["sphere"] == 0
    np.testing.assert_allclose(out[0]["p"], p, atol=1e-6)
    np.testing.assert_allclose(out[0]["n"], -np.array(n) if flip else n, atol=1e-6)
    assert abs(out[0]["depth"] - d) < 1e-6


def test_far_sphere_and_empty():
    t, m = _tree(_unit_tet())
    assert O.single_traversal(t, m, [[3, 3, 3, 0.5]]).shape[0] == 0
    assert O.single_traversal(t, m, np.zeros((0, 4), F32)).shape[0] == 0


def test_traversal_equals_every_tet_on_its_own():
    rng = np.random.default_rng(9)
    mesh = K.box_mesh(3, size=1.0, jitter=0.1, seed=2)
    t, m = _tree(mesh)
    sp = np.concatenate([rng.uniform(-0.7, 0.7, size=(40, 3)), rng.uniform(0.05, 0.3, size=(40, 1))], 1).astype(F32)
    got = O.single_traversal(t, m, sp)
    assert got.shape[0] > 40
    brute = []
    for ti in range(mesh.tets.shape[0]):
        sub = K.Mesh(mesh.X, mesh.Y, mesh.Z, mesh.tets[ti:ti + 1].copy(), mesh.surface[ti:ti + 1].copy())
        ts, ms = _tree(sub)
        b = O.single_traversal(ts, ms, sp)
        b["tet"] = ti
        brute.append(b)
    brute = np.concatenate(brute)
    key = lambda a: np.sort(a, order=["sphere", "tet"]).tobytes()
    assert key(got) == key(brute)
    # grouped by sphere, in sphere order
    assert np.all(np.diff(got["sphere"].astype(np.int64)) >= 0)
