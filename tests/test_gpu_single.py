"""kdop_single_traversal against the oracle's kdop::single_traversal, through
the C ABI: the contact records (sphere, tet, point, normal, depth) bit for bit
and in the same order, for several K, chunk widths, posed bodies, radii and
both normal directions."""
import numpy as np
import pytest

import kdop_amd as K
from kdop_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.mark.parametrize("k,n,chunk_bytes,nsph,flip", [(8, 10, 8000, 4096, False), (26, 12, 8000, 4096, True),
                                                       (14, 12, 8000, 2048, False), (26, 12, 2000, 2048, False),
                                                       (26, 40, 8000, 512, False), (6, 12, 8000, 2048, True),
                                                       (18, 12, 8000, 2048, False), (12, 11, 8000, 1024, False),
                                                       (20, 11, 8000, 1024, True), (8, 10, 32000, 2048, False),
                                                       (26, 10, 2 ** 64 - 1, 1024, True)],
                         ids=["k8_box10", "k26_box12_flip", "k14_box12", "k26_L2", "k26_box40", "k6_box12_flip", "k18_box12",
                              "k12_box11", "k20_box11_flip", "k8_L128", "k26_one_chunk_flip"])
def test_single_traversal_matches_oracle(k, n, chunk_bytes, nsph, flip):
    rng = np.random.default_rng(n * 13 + k)
    mesh = K.box_mesh(n, size=1.0, jitter=0.1, seed=4)
    pose = W.Pose(np.array([-0.2, 0.1, 0.3], F32), K.rng_quaternions(6, 2, 1)[0])
    X, Y, Z = O.transform_f32(mesh.X, mesh.Y, mesh.Z, pose.p, pose.q)
    tree = O.Tree(k, chunk_bytes, mesh.tets, mesh.X, mesh.Y, mesh.Z)
    tree.refit(X, Y, Z)
    ctx = K.Context(K=k, chunk_bytes=chunk_bytes)
    b = ctx.make_tree(mesh)
    ctx.refit([b], pose.p[None, :], pose.q[None, :])
    # centres around the posed box (near its surface and inside), radii from small to large
    c = pose.p[None, :] + rng.uniform(-0.75, 0.75, size=(nsph, 3))
    r = rng.uniform(0.01, 0.15, size=(nsph, 1))
    sp = np.concatenate([c, r], 1).astype(F32)
    got = ctx.single_traversal(b, sp, flip=flip)
    want = O.single_traversal(tree, (mesh.tets, X, Y, Z, mesh.surface), sp, flip=flip)
    print(k, n, chunk_bytes, "contacts", want.shape[0])
    assert want.shape[0] > nsph // 4
    assert got.tobytes() == want.tobytes()


def test_single_traversal_in_batches_is_unchanged():
    """A small frontier limit splits the spheres into batches (halving until
    each fits): the same records; a limit no single sphere fits is an error."""
    k, n = 26, 16
    rng = np.random.default_rng(5)
    mesh = K.box_mesh(n, size=1.0, jitter=0.1, seed=2)
    ctx = K.Context(K=k)
    b = ctx.make_tree(mesh)
    ctx.refit([b], np.zeros((1, 3), F32), np.array([[0, 0, 0, 1]], F32))
    c = rng.uniform(-0.7, 0.7, size=(3000, 3))
    r = rng.uniform(0.01, 0.2, size=(3000, 1))
    sp = np.concatenate([c, r], 1).astype(F32)
    full = ctx.single_traversal(b, sp)
    X, Y, Z = mesh.X, mesh.Y, mesh.Z
    tree = O.Tree(k, 8000, mesh.tets, X, Y, Z)
    tree.refit(X, Y, Z)
    want = O.single_traversal(tree, (mesh.tets, X, Y, Z, mesh.surface), sp)
    assert full.tobytes() == want.tobytes()
    for limit in (20000, 5000):
        ctx.set_option("single_items", limit)
        got = ctx.single_traversal(b, sp)
        print("limit", limit, "contacts", got.shape[0])
        assert got.tobytes() == full.tobytes()
    ctx.set_option("single_items", 1)
    with pytest.raises(K.KdopError):
        ctx.single_traversal(b, sp[:4])
    ctx.set_option("single_items", 0)
    assert ctx.get_option("single_items") == 1 << 28


def test_single_traversal_capacity_and_errors():
    ctx = K.Context(K=8)
    b = ctx.make_tree(K.box_mesh(3))
    sp = np.array([[0, 0, 0.5, 0.2], [5, 5, 5, 0.1]], F32)
    n = ctypes_count(ctx, b, sp)
    assert n >= 1
    with pytest.raises(K.KdopError):
        ctx.single_traversal(b + 9, sp)


def ctypes_count(ctx, b, sp):
    import ctypes
    cnt = ctypes.c_uint64()
    ctx._check(ctx._lib.kdop_single_traversal(ctx._h, b, sp.shape[0], K.ptr(sp), 0, None, 0, ctypes.byref(cnt), 0))
    out = np.zeros(max(cnt.value - 1, 0), K._abi.SPHERE_CONTACT_DTYPE)
    if cnt.value:  # one record short: a capacity error, nothing written
        rc = ctx._lib.kdop_single_traversal(ctx._h, b, sp.shape[0], K.ptr(sp), 0, K.ptr(out) if out.size else None,
                                            out.shape[0], ctypes.byref(cnt), 0)
        assert rc == K._abi.KDOP_ERR_CAPACITY if out.size else rc == 0
    return cnt.value


@pytest.mark.parametrize("k,n", [(8, 1), (26, 1), (6, 2)], ids=["k8_cube", "k26_cube", "k6_2cubes"])
def test_single_traversal_one_chunk_bodies(k, n):
    """Bodies of one branch chunk (no super level, no breadth-first step): the
    same contacts as the oracle, including the single tetrahedron."""
    rng = np.random.default_rng(k * 5 + n)
    tet = K.Mesh(F32([0, 1, 0, 0]), F32([0, 0, 1, 0]), F32([0, 0, 0, 1]), np.array([[0, 1, 2, 3]], np.uint32), np.array([15], np.uint8))
    for mesh in (tet, K.box_mesh(n, size=1.0, jitter=0.0, seed=1)):
        tree = O.Tree(k, 8000, mesh.tets, mesh.X, mesh.Y, mesh.Z)
        tree.refit(mesh.X, mesh.Y, mesh.Z)
        ctx = K.Context(K=k)
        b = ctx.make_tree(mesh)
        ctx.refit([b], np.zeros((1, 3), F32), np.array([[0, 0, 0, 1]], F32))
        assert ctx.shape(b)["H"] == 1
        c = rng.uniform(-0.2, 0.8, size=(512, 3))
        r = rng.uniform(0.02, 0.3, size=(512, 1))
        sp = np.concatenate([c, r], 1).astype(F32)
        got = ctx.single_traversal(b, sp)
        want = O.single_traversal(tree, (mesh.tets, mesh.X, mesh.Y, mesh.Z, mesh.surface), sp)
        assert want.shape[0] > 0
        assert got.tobytes() == want.tobytes()


def test_single_traversal_device_arrays_match_host_arrays():
    """is_device = 1: spheres read in place and records written to device
    memory give the host-array call's records byte for byte and in the same
    order; a device buffer one record short is a capacity error that reports
    the count, and misaligned device spheres are an argument error."""
    import ctypes
    import torch
    rng = np.random.default_rng(77)
    mesh = K.box_mesh(12, size=1.0, jitter=0.1, seed=4)
    pose = W.Pose(np.array([-0.2, 0.1, 0.3], F32), K.rng_quaternions(6, 2, 1)[0])
    ctx = K.Context(K=14)
    b = ctx.make_tree(mesh)
    ctx.refit([b], pose.p[None, :], pose.q[None, :])
    c = pose.p[None, :] + rng.uniform(-0.75, 0.75, size=(3000, 3))
    r = rng.uniform(0.01, 0.15, size=(3000, 1))
    sp = np.concatenate([c, r], 1).astype(F32)
    for flip in (False, True):
        want = ctx.single_traversal(b, sp, flip=flip)
        got = ctx.single_traversal(b, torch.from_numpy(sp).cuda(), flip=flip)
        print("device single traversal flip", flip, "contacts", want.shape[0], flush=True)
        assert got.is_cuda and want.shape[0] > 1000
        assert got.cpu().numpy().tobytes() == want.tobytes()
    dsp = torch.from_numpy(sp).cuda()
    rec = K._abi.SPHERE_CONTACT_DTYPE.itemsize
    n = ctypes.c_uint64()
    short = torch.zeros((want.shape[0] - 1, rec), dtype=torch.uint8, device="cuda")
    rc = ctx._lib.kdop_single_traversal(ctx._h, b, sp.shape[0], dsp.data_ptr(), 1, short.data_ptr(), short.shape[0],
                                        ctypes.byref(n), 1)
    assert rc == K._abi.KDOP_ERR_CAPACITY and n.value == want.shape[0]
    assert int(short.sum()) == 0  # nothing written
    raw = torch.zeros(64, dtype=torch.float32, device="cuda")
    rc = ctx._lib.kdop_single_traversal(ctx._h, b, 2, raw.data_ptr() + 4, 0, short.data_ptr(), short.shape[0], ctypes.byref(n), 1)
    assert rc == K._abi.KDOP_ERR_ARG
    ctx.close()
