"""PROX contact reduction on the device (kdop_reduce_contacts in place on a
traversal's contacts, kdop_reduce_prox_contacts on caller arrays) against the
oracle's std::sort + std::unique restatement (prox_collision_detection.h:
340-345), through the C ABI.

Bar: exact -- the same records, bit for bit, in the same order, which also
fixes which of a group of tolerance-equal contacts survives."""
import ctypes

import numpy as np
import pytest
import torch

import kdop_amd as K
from kdop_amd import _abi
from kdop_amd import workloads as W
from oracle import oracle as O
from helpers import device_run
from prox_cases import DTYPE, cases

pytestmark = pytest.mark.gpu

assert DTYPE == _abi.PROX_CONTACT_DTYPE


def _oracle(a):
    return O.prox_reduce(a["body_i"], a["body_j"], a["p"], a["n"], a["depth"])


@pytest.fixture(scope="module")
def ctx():
    c = K.Context(K=8)
    yield c
    c.close()


def test_known_answers(ctx):
    for name, recs, kept in cases():
        got = ctx.reduce_prox_contacts(recs)
        print(name, got.shape[0], flush=True)
        assert got.tobytes() == recs[kept].tobytes(), name


def test_empty_single_and_alias(ctx):
    assert ctx.reduce_prox_contacts(np.zeros(0, DTYPE)).shape[0] == 0
    one = cases()[-1][1]
    assert ctx.reduce_prox_contacts(one).tobytes() == one.tobytes()
    n = ctypes.c_uint64()
    rc = ctx._lib.kdop_reduce_prox_contacts(ctx._h, 1, _abi.ptr(one), _abi.ptr(one), 0, ctypes.byref(n))
    assert rc == _abi.KDOP_ERR_ARG


def _chains(n, seed, span):
    """Contacts packed within a few ulps of each other, so that equal-within-
    precision runs and the last-kept rule matter: p.y = 1 + k ulps with k in
    [0, span), every other float one of a few values a few ulps apart."""
    rng = np.random.default_rng(seed)
    ulp = np.float32(2.0 ** -23)
    near = lambda shape, m: np.float32(1.0) + rng.integers(0, m, shape).astype(np.float32) * ulp
    a = np.zeros(n, DTYPE)
    a["body_i"] = rng.integers(0, 2, n)
    a["body_j"] = rng.integers(0, 2, n)
    a["p"][:, 0] = near(n, 24)
    a["p"][:, 1] = near(n, span)
    a["p"][:, 2] = near(n, 12)
    a["n"] = near((n, 3), 12)
    a["depth"] = near(n, 30)
    return a


@pytest.mark.parametrize("n,span,seed", [(5000, 64, 1), (300000, 200000, 2), (1500000, 1 << 22, 3)])
def test_chains_against_oracle(ctx, n, span, seed):
    a = _chains(n, seed, span)
    want = a[_oracle(a)]
    got = ctx.reduce_prox_contacts(a)
    print("n", n, "kept", got.shape[0], flush=True)
    assert 0 < got.shape[0] < n
    assert got.tobytes() == want.tobytes()


def test_device_pointers(ctx):
    a = _chains(40000, 7, 4096)
    want = a[_oracle(a)]
    src = torch.from_numpy(a.view(np.uint8).copy()).cuda()
    dst = torch.zeros_like(src)
    n = ctypes.c_uint64()
    rc = ctx._lib.kdop_reduce_prox_contacts(ctx._h, a.shape[0], src.data_ptr(), dst.data_ptr(), 1, ctypes.byref(n))
    assert rc == 0
    got = dst.cpu().numpy()[: n.value * DTYPE.itemsize].view(DTYPE)
    assert got.tobytes() == want.tobytes()


def test_reduce_traversal_contacts():
    wl = W.batched_pairs(npairs=12, K=8, seed=5, tmin=1000, tmax=20000)
    pairs = np.concatenate([wl.pairs, wl.pairs[:4], wl.pairs[:4, ::-1]]).astype(np.uint32)
    wl2 = W.PairWorkload(wl.name, wl.K, wl.meshes, pairs, wl.poses)
    dev = device_run(wl2)
    ctx = dev["ctx"]
    c0 = dev["contacts"]
    ids = np.array(dev["ids"], np.uint32)
    bp = ids[pairs]
    a = np.zeros(c0.shape[0], DTYPE)
    a["body_i"] = bp[c0["pair"], 0]
    a["body_j"] = bp[c0["pair"], 1]
    a["p"], a["n"], a["depth"] = c0["p"], c0["n"], c0["depth"]
    keep = _oracle(a)
    dup = int(np.isin(c0["pair"], np.arange(12, 16)).sum())  # contacts of the repeated pairs
    n = ctx.reduce_contacts()
    got = ctx.contacts()
    print("contacts", c0.shape[0], "kept", n, "repeated-pair contacts", dup, flush=True)
    assert dup > 0
    assert n == got.shape[0] == keep.shape[0] <= c0.shape[0] - dup
    assert got.tobytes() == c0[keep].tobytes()
    # the reduced list sorts like any other, and a new traversal brings back the full list
    ctx.sort_contacts("sequential")
    s = ctx.contacts()
    rec = lambda x: sorted(x[i:i + 1].tobytes() for i in range(x.shape[0]))
    assert rec(s) == rec(got)
    assert ctx.reduce_contacts() == n  # reducing a reduced list changes nothing
    ctx.tandem_traversal(np.array([[ids[a_], ids[b_]] for a_, b_ in pairs], np.uint32), narrow=True)
    assert ctx.contacts().shape[0] == c0.shape[0]
    ctx.close()
