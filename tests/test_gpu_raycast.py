"""kdop_raycast against the oracle's kdop::raycast, through the C ABI: hit
flags, lengths and hit points bit for bit, for several K, chunk widths (the
super levels the device culls with), posed bodies and rays that miss, hit,
graze and start inside."""
import numpy as np
import pytest

import kdop_amd as K
from kdop_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32 = np.float32


def _rays(rng, n, center, radius, spread):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = center + radius * u
    target = center + rng.uniform(-spread, spread, size=(n, 3))
    d = target - o
    # a few rays from inside the body and a few pointing away
    o[: n // 16] = center + rng.uniform(-0.1, 0.1, size=(n // 16, 3))
    d[n // 16: n // 8] = -d[n // 16: n // 8]
    return o.astype(F32), d.astype(F32)


@pytest.mark.parametrize("k,n,chunk_bytes,nrays", [(8, 10, 8000, 4096), (26, 12, 8000, 4096), (14, 12, 8000, 2048),
                                                   (26, 12, 2000, 2048), (26, 40, 8000, 256), (6, 12, 8000, 2048),
                                                   (18, 12, 8000, 2048), (12, 11, 8000, 1024), (20, 11, 8000, 1024),
                                                   (8, 10, 32000, 2048), (14, 10, 2 ** 64 - 1, 1024)],
                         ids=["k8_box10", "k26_box12", "k14_box12", "k26_L2", "k26_box40", "k6_box12", "k18_box12", "k12_box11",
                              "k20_box11", "k8_L128", "k14_one_chunk"])
def test_raycast_matches_oracle(k, n, chunk_bytes, nrays):
    rng = np.random.default_rng(n * 31 + k)
    mesh = K.box_mesh(n, size=1.0, jitter=0.1, seed=7)
    pose = W.Pose(np.array([0.3, -0.2, 0.1], F32), K.rng_quaternions(5, 1, 1)[0])
    X, Y, Z = O.transform_f32(mesh.X, mesh.Y, mesh.Z, pose.p, pose.q)
    tree = O.Tree(k, chunk_bytes, mesh.tets, mesh.X, mesh.Y, mesh.Z)
    tree.refit(X, Y, Z)
    ctx = K.Context(K=k, chunk_bytes=chunk_bytes)
    b = ctx.make_tree(mesh)
    ctx.refit([b], pose.p[None, :], pose.q[None, :])
    o, d = _rays(rng, nrays, pose.p.astype(np.float64), 2.0, 0.55)
    h_d, l_d, f_d = ctx.raycast(b, o, d)
    h_o, l_o, f_o = O.raycast(tree, (mesh.tets, X, Y, Z, mesh.surface), o, d)
    print(k, n, chunk_bytes, "hits", int(f_o.sum()), "of", nrays)
    assert f_o.sum() > nrays // 4
    np.testing.assert_array_equal(f_d, f_o)
    assert l_d.tobytes() == l_o.tobytes()
    assert h_d.tobytes() == h_o.tobytes()


def test_raycast_errors_and_empty():
    ctx = K.Context(K=8)
    b = ctx.make_tree(K.box_mesh(2))
    h, l, f = ctx.raycast(b, np.zeros((0, 3), F32), np.zeros((0, 3), F32))
    assert h.shape == (0, 3)
    with pytest.raises(K.KdopError):
        ctx.raycast(b + 5, np.zeros((1, 3), F32), np.ones((1, 3), F32))


def _axis_rays(rng, n, lo, hi):
    """Rays along the coordinate axes and the face and cube diagonals (parallel
    to many slabs of the direction tables, the f == 0 branch), aimed at points
    of a grid of spacing (hi - lo) / 8 (along faces and edges of an unjittered
    cube grid when the body is unposed), from outside the box."""
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (0, -1, 1), (1, 0, -1), (1, 1, 1),
            (-1, 1, -1)]
    d = np.array([dirs[i % len(dirs)] for i in range(n)], np.float64)
    g = lo + (hi - lo) * rng.integers(0, 9, size=(n, 3)) / 8.0
    o = g - 3.0 * d / np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(F32), d.astype(F32)


@pytest.mark.parametrize("k,jitter,posed", [(26, 0.0, False), (26, 0.1, False), (14, 0.0, True), (8, 0.0, False), (6, 0.1, True)],
                         ids=["k26_grid", "k26_jitter", "k14_posed", "k8_grid", "k6_posed"])
def test_raycast_axis_and_diagonal_rays(k, jitter, posed):
    rng = np.random.default_rng(k + (7 if posed else 0))
    mesh = K.box_mesh(10, size=1.0, jitter=jitter, seed=3)
    pose = W.Pose(np.array([0.05, 0.0, -0.1], F32), K.rng_quaternions(9, 0, 1)[0]) if posed else W.IDENTITY
    X, Y, Z = O.transform_f32(mesh.X, mesh.Y, mesh.Z, pose.p, pose.q)
    tree = O.Tree(k, 8000, mesh.tets, mesh.X, mesh.Y, mesh.Z)
    tree.refit(X, Y, Z)
    ctx = K.Context(K=k)
    b = ctx.make_tree(mesh)
    ctx.refit([b], pose.p[None, :], pose.q[None, :])
    lo = np.array([X.min(), Y.min(), Z.min()], np.float64)
    hi = np.array([X.max(), Y.max(), Z.max()], np.float64)
    o, d = _axis_rays(rng, 3000, lo, hi)
    h_d, l_d, f_d = ctx.raycast(b, o, d)
    h_o, l_o, f_o = O.raycast(tree, (mesh.tets, X, Y, Z, mesh.surface), o, d)
    print(k, jitter, posed, "hits", int(f_o.sum()), "of", o.shape[0])
    assert f_o.sum() > o.shape[0] // 4
    np.testing.assert_array_equal(f_d, f_o)
    assert l_d.tobytes() == l_o.tobytes()
    assert h_d.tobytes() == h_o.tobytes()


def test_raycast_in_batches_is_unchanged():
    """A small frontier limit splits the rays into batches: the same results."""
    rng = np.random.default_rng(3)
    mesh = K.box_mesh(16, size=1.0, jitter=0.1, seed=5)
    ctx = K.Context(K=26)
    b = ctx.make_tree(mesh)
    ctx.refit([b], np.zeros((1, 3), F32), np.array([[0, 0, 0, 1]], F32))
    o, d = _rays(rng, 3000, np.zeros(3), 2.0, 0.55)
    full = ctx.raycast(b, o, d)
    ctx.set_option("single_items", 20000)
    part = ctx.raycast(b, o, d)
    for a, c in zip(full, part):
        assert a.tobytes() == c.tobytes()
    ctx.set_option("single_items", 1)
    with pytest.raises(K.KdopError):  # rays from outside, aimed at the box (the first ones start inside it)
        ctx.raycast(b, o[-4:], d[-4:])


@pytest.mark.parametrize("k,n", [(8, 1), (26, 1), (6, 2)], ids=["k8_cube", "k26_cube", "k6_2cubes"])
def test_raycast_one_chunk_bodies(k, n):
    """Bodies of one branch chunk (no super level, no breadth-first step): the
    same hits as the oracle, including the single tetrahedron."""
    rng = np.random.default_rng(k * 7 + n)
    tet = K.Mesh(np.array([0, 1, 0, 0], F32), np.array([0, 0, 1, 0], F32), np.array([0, 0, 0, 1], F32),
                 np.array([[0, 1, 2, 3]], np.uint32), np.array([15], np.uint8))
    for mesh in (tet, K.box_mesh(n, size=1.0, jitter=0.0, seed=1)):
        tree = O.Tree(k, 8000, mesh.tets, mesh.X, mesh.Y, mesh.Z)
        tree.refit(mesh.X, mesh.Y, mesh.Z)
        ctx = K.Context(K=k)
        b = ctx.make_tree(mesh)
        ctx.refit([b], np.zeros((1, 3), F32), np.array([[0, 0, 0, 1]], F32))
        assert ctx.shape(b)["H"] == 1
        o, d = _rays(rng, 512, np.full(3, 0.25), 2.0, 0.3)
        got = ctx.raycast(b, o, d)
        want = O.raycast(tree, (mesh.tets, mesh.X, mesh.Y, mesh.Z, mesh.surface), o, d)
        assert want[2].sum() > 0
        for a, w in zip(got, want):
            assert np.asarray(a).tobytes() == np.asarray(w).tobytes()


def test_raycast_device_arrays_match_host_arrays():
    """is_device = 1: rays read from and results written to device memory in
    place give the host-array call's results byte for byte (and so the
    oracle's); a misaligned device pointer is an argument error."""
    import torch
    rng = np.random.default_rng(41)
    mesh = K.box_mesh(12, size=1.0, jitter=0.1, seed=7)
    pose = W.Pose(np.array([0.3, -0.2, 0.1], F32), K.rng_quaternions(5, 1, 1)[0])
    ctx = K.Context(K=26)
    b = ctx.make_tree(mesh)
    ctx.refit([b], pose.p[None, :], pose.q[None, :])
    o, d = _rays(rng, 3000, pose.p.astype(np.float64), 2.0, 0.55)
    h, l, f = ctx.raycast(b, o, d)
    hd, ld, fd = ctx.raycast(b, torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda())
    print("device raycast hits", int(fd.sum()), "of", o.shape[0], flush=True)
    assert fd.is_cuda and ld.is_cuda and hd.is_cuda
    assert fd.cpu().numpy().astype(bool).tolist() == f.tolist()
    assert ld.cpu().numpy().tobytes() == l.tobytes()
    assert hd.cpu().numpy().tobytes() == h.tobytes()
    # the second call reuses the context's scratch after a host-array call
    hd2, ld2, _ = ctx.raycast(b, torch.from_numpy(o[:700]).cuda(), torch.from_numpy(d[:700]).cuda())
    assert ld2.cpu().numpy().tobytes() == l[:700].tobytes()
    raw = torch.zeros(3 * 8 + 1, dtype=torch.float32, device="cuda")
    import ctypes
    rc = ctx._lib.kdop_raycast(ctx._h, b, 1, raw.data_ptr() + 2, raw.data_ptr(), raw.data_ptr() + 16, raw.data_ptr() + 64,
                               raw.data_ptr() + 80, 1)
    assert rc == K._abi.KDOP_ERR_ARG
    ctx.close()
