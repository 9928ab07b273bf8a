"""Adjoint of the penetration measure (kdop_narrow_adjoint) and its
torch.autograd harness, against the FP64 oracle gradient and central finite
differences (h = 1e-3).

Phi(pair) = sum of the depths of the pair's reported contacts, with the
contact set and every discrete choice of the traversal frozen.  Bars: Phi and
the gradient within 1e-5 (relative to the largest component) of the oracle's
FP64 restatement with the same frozen choices (the device computes in FP64 and
rounds its outputs to float32); the gradient within 1e-4 of central
differences of Phi.
"""
import numpy as np
import pytest
import torch

import kdop_amd as K
from kdop_amd import workloads as W
from kdop_amd.autograd import penetration_measure
from helpers import narrow_records, oracle_run
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ORACLE_RTOL = 1e-5
FD_RTOL = 1e-4
H = 1e-3


def _context(wl):
    ctx = K.Context(K=wl.K)
    ids = [ctx.make_tree(m) for m in wl.meshes]
    pos = np.stack([p.p for p in wl.poses]).astype(np.float32)
    quat = np.stack([p.q for p in wl.poses]).astype(np.float32)
    pairs = np.array([[ids[a], ids[b]] for a, b in wl.pairs], np.uint32)
    return ctx, np.array(ids, np.uint32), pos, quat, pairs


def _records(recs):
    R = np.zeros(len(recs), dtype=[("ta", np.uint32), ("tb", np.uint32), ("face", np.uint8), ("nuniq", np.uint8),
                                   ("cmin", np.uint8), ("cmax", np.uint8)])
    for i, (ta_i, tb_i, r) in enumerate(recs):
        R[i] = (ta_i, tb_i, r["face"], r["pts"].shape[0], r["codes"][r["imin"]], r["codes"][r["imax"]])
    return R


def test_adjoint_matches_fp64_oracle():
    """One jittered pair (K=14): device Phi and pose gradients of both bodies
    vs the oracle's FP64 frozen evaluation and reverse mode."""
    wl = W.large_pair(n=8, K=14, seed=17, jitter=0.1)
    ctx, ids, pos, quat, pairs = _context(wl)
    ctx.refit(ids, pos, quat)
    c = ctx.tandem_traversal(pairs, narrow=True)
    assert c["contacts"] > 10
    phi, gp, gq = ctx.narrow_adjoint(ids)
    recs, _, _, ma, mb = narrow_records(wl)
    R = _records(recs)
    to64 = lambda v: np.asarray(v, np.float64)  # noqa: E731  (the float32 poses the device used)
    want = O.phi_frozen_f64(R, ma, (to64(pos[0]), to64(quat[0])), mb, (to64(pos[1]), to64(quat[1])), grad=True)
    phi_o, g_o = want[0], np.concatenate(want[1:])
    g_d = np.concatenate([gp[0].cpu().numpy(), gq[0].cpu().numpy(), gp[1].cpu().numpy(), gq[1].cpu().numpy()])
    assert abs(float(phi[0]) - phi_o) <= ORACLE_RTOL * max(1.0, abs(phi_o))
    # Phi equals the sum of the reported contact depths (float32 forward)
    depths = ctx.contacts()["depth"].astype(np.float64).sum()
    assert abs(depths - phi_o) <= 1e-4 * max(1.0, abs(phi_o))
    scale = max(np.abs(g_o).max(), 1e-6)
    assert np.abs(g_d - g_o).max() <= ORACLE_RTOL * scale, (g_d, g_o)


def _fd_check(ctx, ids, pos, quat, pairs, which, comps):
    """Central differences of the device's frozen Phi (refit without
    re-traversal) vs its analytic gradient, for bodies `which`."""
    phi0, gp, gq = ctx.narrow_adjoint(ids)
    g = np.concatenate([gp.cpu().numpy(), gq.cpu().numpy()], axis=1)  # (n, 7)
    worst, scale = 0.0, 0.0
    for b in which:
        for j in comps:
            vals, xs = [], []
            for sgn in (1.0, -1.0):
                P, Q = pos.copy(), quat.copy()
                if j < 3:
                    P[b, j] += np.float32(sgn * H)
                    xs.append(float(P[b, j]))
                else:
                    Q[b, j - 3] += np.float32(sgn * H)
                    xs.append(float(Q[b, j - 3]))
                ctx.refit(ids, P, Q)
                vals.append(float(ctx.narrow_adjoint(ids)[0].double().sum()))
            fd = (vals[0] - vals[1]) / (xs[0] - xs[1])  # the float32 step actually taken
            worst = max(worst, abs(fd - g[b, j]))
            scale = max(scale, abs(fd), abs(g[b, j]))
    ctx.refit(ids, pos, quat)
    return worst, scale


def test_adjoint_matches_central_differences_batched():
    """Four overlapping pairs of ~1-2k-tet bodies (K=14) in one traversal:
    the gradient of sum(Phi) w.r.t. every pose component of every body vs
    central differences, h = 1e-3.  The perturbed evaluations reuse the
    forward's frozen choices (refit, no re-traversal), as the FP64 oracle's
    finite differences do."""
    wl = W.batched_pairs(npairs=4, K=14, seed=5, tmin=1000, tmax=2000, separated=0.0)
    ctx, ids, pos, quat, pairs = _context(wl)
    ctx.refit(ids, pos, quat)
    c = ctx.tandem_traversal(pairs, narrow=True)
    assert c["contacts"] > 10
    worst, scale = _fd_check(ctx, ids, pos, quat, pairs, range(len(ids)), range(7))
    assert worst <= FD_RTOL * scale, (worst, scale)


def test_autograd_function_backward_is_the_adjoint():
    """PenetrationMeasure: forward Phi equals narrow_adjoint's Phi, and
    backward of sum(w * Phi) equals the adjoint seeded with w."""
    wl = W.batched_pairs(npairs=3, K=14, seed=8, tmin=1000, tmax=3000, separated=0.0)
    ctx, ids, pos, quat, pairs = _context(wl)
    p = torch.tensor(pos, device="cuda", requires_grad=True)
    q = torch.tensor(quat, device="cuda", requires_grad=True)
    phi = penetration_measure(ctx, ids, pairs, p, q)
    w = torch.tensor([1.0, -2.0, 0.5], device="cuda")
    (phi * w).sum().backward()
    phi_ref, gp, gq = ctx.narrow_adjoint(ids, dphi=w)
    torch.testing.assert_close(phi, phi_ref, rtol=0, atol=0)
    torch.testing.assert_close(p.grad, gp, rtol=0, atol=0)
    torch.testing.assert_close(q.grad, gq, rtol=0, atol=0)
    # the forward's Phi is the sum of the contact depths of each pair
    con = ctx.contacts()
    for k in range(pairs.shape[0]):
        d = con["depth"][con["pair"] == k].astype(np.float64).sum()
        assert abs(d - float(phi[k].detach())) <= 1e-4 * max(1.0, abs(d))


def test_autograd_rejects_a_stale_context():
    wl = W.batched_pairs(npairs=2, K=14, seed=9, tmin=1000, tmax=1500, separated=0.0)
    ctx, ids, pos, quat, pairs = _context(wl)
    p = torch.tensor(pos, device="cuda", requires_grad=True)
    q = torch.tensor(quat, device="cuda", requires_grad=True)
    phi = penetration_measure(ctx, ids, pairs, p, q)
    ctx.tandem_traversal(pairs, narrow=True)
    with pytest.raises(RuntimeError):
        phi.sum().backward()


def test_adjoint_c5_scale():
    """64 overlapping pairs of ~53k-tet jittered boxes (K=14, 30-60 % AABB
    overlap): every body gets a finite gradient, Phi matches the contact
    depths per pair, and central differences agree on a sample of bodies."""
    wl = W.batched_pairs(npairs=64, K=14, seed=29, tmin=5e4, tmax=5e4, separated=0.0)
    ctx, ids, pos, quat, pairs = _context(wl)
    ctx.refit(ids, pos, quat)
    c = ctx.tandem_traversal(pairs, narrow=True)
    assert c["narrow_pairs"] > 1000
    phi, gp, gq = ctx.narrow_adjoint(ids)
    assert torch.isfinite(gp).all() and torch.isfinite(gq).all()
    con = ctx.contacts()
    d = np.bincount(con["pair"], weights=con["depth"].astype(np.float64), minlength=pairs.shape[0])
    np.testing.assert_allclose(phi.cpu().numpy(), d, rtol=1e-4, atol=1e-4)
    worst, scale = _fd_check(ctx, ids, pos, quat, pairs, [0, 1, 77], range(7))
    assert worst <= FD_RTOL * scale, (worst, scale)


def test_adjoint_c5_every_body_vs_fp64_oracle():
    """The C5 draw (64 overlapping pairs of ~53k-tet jittered boxes, K=14):
    Phi of every pair and the pose gradients of all 128 bodies against the
    oracle's FP64 frozen evaluation and reverse mode, on the records the
    oracle's float32 narrow phase chooses for each pair's surface pairs; the
    same bar as the single pair."""
    wl = W.batched_pairs(npairs=64, K=14, seed=29, tmin=5e4, tmax=5e4, separated=0.0)
    ctx, ids, pos, quat, pairs = _context(wl)
    ctx.refit(ids, pos, quat)
    ctx.tandem_traversal(pairs, narrow=True)
    phi, gp, gq = (t.cpu().numpy() for t in ctx.narrow_adjoint(ids))
    orc, _, worlds = oracle_run(wl, narrow=False, threads=8)
    to64 = lambda v: np.asarray(v, np.float64)  # noqa: E731
    checked = 0
    for i, (a, b) in enumerate(wl.pairs):
        ma, mb = wl.meshes[a], wl.meshes[b]
        PA, PB = np.stack(worlds[a], 1), np.stack(worlds[b], 1)
        recs = []
        for ta_i, tb_i in orc[i][2]:
            r = O.contacts_pair_codes(PA[ma.tets[ta_i]], PB[mb.tets[tb_i]], ma.surface[ta_i], mb.surface[tb_i])
            if r["pts"].shape[0]:
                recs.append((int(ta_i), int(tb_i), r))
        want = O.phi_frozen_f64(_records(recs), ma, (to64(pos[a]), to64(quat[a])), mb, (to64(pos[b]), to64(quat[b])),
                                grad=True)
        g_o = np.concatenate(want[1:])
        g_d = np.concatenate([gp[a], gq[a], gp[b], gq[b]])
        assert abs(float(phi[i]) - want[0]) <= ORACLE_RTOL * max(1.0, abs(want[0])), (i, float(phi[i]), want[0])
        assert np.abs(g_d - g_o).max() <= ORACLE_RTOL * max(np.abs(g_o).max(), 1e-6), (i, g_d, g_o)
        checked += len(recs) > 0
    assert checked >= 60
