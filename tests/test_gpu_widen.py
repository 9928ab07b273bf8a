"""Device contact sort (kdop_sort_contacts) and device Morton reorder
(kdop_mesh_reorder_gpu) against the oracle, through the C ABI.

Bar: exact.  KDOP_ORDER_SEQUENTIAL must give the oracle's own report order
(the sequential reference's callback order, contact for contact, bit for
bit); KDOP_ORDER_OPENCL must give numpy's stable lexsort by (pair, p, n,
depth) of the array it sorted; the reorder must equal the oracle's byte for
byte."""
import numpy as np
import pytest

import kdop_amd as K
from kdop_amd import workloads as W
from oracle import oracle as O
from helpers import device_run, oracle_contacts_with_ordinals, oracle_run

pytestmark = pytest.mark.gpu


def _oracle_report(orc):
    """The oracle's contacts of every pair, in its report order, as device records."""
    rows = []
    for pi, (_, _, _, con) in enumerate(orc):
        ords = oracle_contacts_with_ordinals(con)
        for c, o in zip(con, ords):
            rows.append((pi, c["ta"], c["tb"], o, c["p"], c["n"], c["d"]))
    out = np.zeros(len(rows), K.CONTACT_DTYPE)
    for i, (pi, ta, tb, o, p, n, d) in enumerate(rows):
        out[i] = (pi, ta, tb, o, p, n, d)
    return out


@pytest.mark.parametrize("wl_name", ["batched14", "pair26", "pair8_L128", "pair14_one_chunk"])
def test_sort_contacts_sequential_is_the_report_order(wl_name):
    """Also for chunks of 128 leaves and one chunk of every tet (mem_bytes
    SIZE_MAX), whose descent paths run 7 and 13 steps deep."""
    chunk_bytes = {"pair8_L128": 32000, "pair14_one_chunk": 2 ** 64 - 1}.get(wl_name, 8000)
    if wl_name == "batched14":
        wl = W.batched_pairs(npairs=16, K=14, seed=4, tmin=500, tmax=8000)
    elif wl_name == "pair8_L128":
        wl = W.large_pair(n=10, K=8, seed=3)
    elif wl_name == "pair14_one_chunk":
        wl = W.large_pair(n=10, K=14, seed=3)
    else:
        wl = W.large_pair(n=12, K=26, seed=3)
    orc, _, _ = oracle_run(wl, chunk_bytes=chunk_bytes)
    dev = device_run(wl, chunk_bytes=chunk_bytes)
    ctx = dev["ctx"]
    ctx.sort_contacts("sequential")
    got = ctx.contacts()
    want = _oracle_report(orc)
    print(wl_name, "contacts", got.shape[0])
    assert got.shape[0] == want.shape[0] > 0
    if wl_name == "batched14":
        assert got.shape[0] > 1024  # more than one sort tile
    assert got.tobytes() == want.tobytes()


def test_sort_contacts_opencl_order_and_idempotence():
    wl = W.batched_pairs(npairs=32, K=8, seed=9, tmin=1000, tmax=20000)
    ctx = device_run(wl, narrow=True, keep_all=False)["ctx"]
    c0 = ctx.contacts()
    ctx.sort_contacts("opencl")
    c1 = ctx.contacts()
    o = np.lexsort((c0["depth"], c0["n"][:, 2], c0["n"][:, 1], c0["n"][:, 0], c0["p"][:, 2], c0["p"][:, 1], c0["p"][:, 0], c0["pair"]))
    print("contacts", c0.shape[0])
    assert c0.shape[0] > 4096
    assert c1.tobytes() == c0[o].tobytes()
    ctx.sort_contacts("opencl")  # sorting sorted contacts changes nothing
    assert ctx.contacts().tobytes() == c1.tobytes()
    ctx.sort_contacts("sequential")  # the two orders hold the same contacts
    c2 = ctx.contacts()
    rec = lambda a: sorted(a[i:i + 1].tobytes() for i in range(a.shape[0]))
    assert rec(c2) == rec(c1)


def test_sort_contacts_small_and_errors():
    tet = K.Mesh(np.array([0, 1, 0, 0], np.float32), np.array([0, 0, 1, 0], np.float32), np.array([0, 0, 0, 1], np.float32),
                 np.array([[0, 1, 2, 3]], np.uint32), np.array([15], np.uint8))
    wl = W.PairWorkload("tets", 8, [tet, tet], np.array([[0, 1]], np.uint32),
                        [W.IDENTITY, W.Pose(np.array([0.1, 0.1, 0.1], np.float32), np.array([0, 0, 0, 1], np.float32))])
    dev = device_run(wl)
    ctx = dev["ctx"]
    c0 = ctx.contacts()
    ctx.sort_contacts("sequential")
    assert ctx.contacts().shape[0] == c0.shape[0]
    with pytest.raises(K.KdopError):
        ctx._check(ctx._lib.kdop_sort_contacts(ctx._h, 7))
    fresh = K.Context(K=8)
    with pytest.raises(K.KdopError):
        fresh.sort_contacts("sequential")  # no traversal yet


def _mesh_case(name):
    rng = np.random.default_rng(17)
    if name in ("box12_shuffled", "box40"):
        n = 12 if name == "box12_shuffled" else 40
        X, Y, Z, T = K.box_mesh_raw(n, size=1.0, jitter=0.1, seed=3)
        if name == "box12_shuffled":
            T = T[rng.permutation(T.shape[0])]
            pv = rng.permutation(X.shape[0]).astype(np.uint32)
            inv = np.empty_like(pv)
            inv[pv] = np.arange(pv.shape[0], dtype=np.uint32)
            X, Y, Z, T = X[pv], Y[pv], Z[pv], inv[T]
        return X, Y, Z, T
    if name == "negative":
        X, Y, Z, T = K.box_mesh_raw(6, size=1.0, jitter=0.1, seed=4)
        return X - np.float32(10), Y - np.float32(12), Z - np.float32(7), T
    if name == "ties":  # the same tet repeated: equal centroids, equal codes
        X, Y, Z, T = K.box_mesh_raw(4, size=1.0, jitter=0.0, seed=0)
        T = np.concatenate([T, T[::-1], T[::3]])
        return X, Y, Z, T[rng.permutation(T.shape[0])]
    if name == "flat":  # one axis without extent
        X, Y, Z, T = K.box_mesh_raw(5, size=1.0, jitter=0.1, seed=6)
        return X, Y, np.zeros_like(Z), T
    if name == "unused_verts":
        X, Y, Z, T = K.box_mesh_raw(5, size=1.0, jitter=0.1, seed=7)
        keep = T[: T.shape[0] // 2]
        return X, Y, Z, keep
    if name == "single_tet":
        return (np.array([0, 1, 0, 0], np.float32), np.array([0, 0, 1, 0], np.float32), np.array([0, 0, 0, 1], np.float32),
                np.array([[3, 1, 2, 0]], np.uint32))
    raise ValueError(name)


@pytest.mark.parametrize("name", ["box12_shuffled", "box40", "negative", "ties", "flat", "unused_verts", "single_tet"])
def test_mesh_reorder_gpu_matches_oracle(name):
    X, Y, Z, T = _mesh_case(name)
    ctx = K.Context(K=8)
    got = ctx.mesh_reorder(X, Y, Z, T)
    want = O.mesh_reorder(X, Y, Z, T)
    for g, w in zip(got, want):
        assert g.tobytes() == w.tobytes()


def test_mesh_reorder_gpu_rejects_bad_indices():
    ctx = K.Context(K=8)
    X = np.zeros(4, np.float32)
    with pytest.raises(K.KdopError):
        ctx.mesh_reorder(X, X, X, np.array([[0, 1, 2, 4]], np.uint32))
