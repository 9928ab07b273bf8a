"""Allocation-free PROX contact reduction (kdop_reduce_prox_contacts_noalloc)
against the allocating entry point (kdop_reduce_prox_contacts on device
arrays), through the C ABI: the same records, bit for bit and in order, on two
calls in a row, the second smaller than the first; a call over the limit is
rejected before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

import kdop_amd as K
from kdop_amd import _abi

pytestmark = pytest.mark.gpu

DT = _abi.PROX_CONTACT_DTYPE


def _contacts(n, seed):
    """n records on a grid a few ulps apart (many equal within the working
    precision, so the reduction drops some), bodies from a small set."""
    rng = np.random.default_rng(seed)
    ulp = np.float32(2.0 ** -23)
    a = np.zeros(n, DT)
    a["body_i"] = rng.integers(0, 4, n)
    a["body_j"] = a["body_i"] + rng.integers(1, 3, n)
    a["p"] = np.float32(1.0) + rng.integers(0, 64, (n, 3)).astype(np.float32) * (4 * ulp)
    a["n"] = np.float32(0.5) + rng.integers(0, 3, (n, 3)).astype(np.float32) * ulp
    a["depth"] = -rng.integers(0, 5, n).astype(np.float32) * np.float32(1e-3)
    return a


def _reduce(ctx, recs, noalloc):
    din = torch.from_numpy(recs.view(np.uint8).copy()).cuda()
    dout = torch.zeros_like(din)
    torch.cuda.synchronize()
    n = ctypes.c_uint64()
    if noalloc:
        rc = ctx._lib.kdop_reduce_prox_contacts_noalloc(ctx._h, recs.shape[0], din.data_ptr(), dout.data_ptr(), ctypes.byref(n))
    else:
        rc = ctx._lib.kdop_reduce_prox_contacts(ctx._h, recs.shape[0], din.data_ptr(), dout.data_ptr(), 1, ctypes.byref(n))
    assert rc == 0, ctx._lib.kdop_last_error(ctx._h).decode()
    return dout.cpu().numpy().view(DT)[:n.value].copy()


def test_noalloc_reduction_matches_the_allocating_entry_point():
    ctx = K.Context(K=8)
    for n, seed in ((_abi.KDOP_NOALLOC_MAX_CONTACTS, 1), (3000, 2)):
        recs = _contacts(n, seed)
        got = _reduce(ctx, recs, True)
        want = _reduce(ctx, recs, False)
        print("contacts", n, "kept", got.shape[0], flush=True)
        assert 0 < got.shape[0] < n
        assert got.tobytes() == want.tobytes()
    # over the limit: rejected before any work (the arrays are never read)
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    kept = ctypes.c_uint64(7)
    rc = ctx._lib.kdop_reduce_prox_contacts_noalloc(ctx._h, _abi.KDOP_NOALLOC_MAX_CONTACTS + 1, buf.data_ptr(),
                                                    buf[32:].data_ptr(), ctypes.byref(kept))
    assert rc == _abi.KDOP_ERR_CAPACITY and kept.value == 0
    assert str(_abi.KDOP_NOALLOC_MAX_CONTACTS) in ctx._lib.kdop_last_error(ctx._h).decode()
    ctx.close()
