"""Seeded synthetic workloads (box-grid tetrahedral meshes + rigid poses).

Every input comes from the library's own generator and seeded stream
(kdop_make_box_mesh, kdop_rng_*), so the CPU oracle and the device see the
same bits on any host of this image.

    single_pair(n=10, K=8)        one pair of unit boxes, body B rotated by a seeded
                                  uniform quaternion and moved by (0.05, 0.05, 0.05)
    large_pair(n=40, K=26)        jittered boxes, pose drawn until the world AABB
                                  overlap fraction is in [0.3, 0.6]
    batched_pairs(npairs=512)     meshes of log-uniform tet counts, ~40 % of the
                                  pairs separated
    self_contact(n=20, K=26)      a jittered box against an identical copy at identity
    motion(...)                   seeded rigid motion sequence for refit series
"""
from dataclasses import dataclass, field

import numpy as np

from . import box_mesh, rng_normal, rng_quaternions, rng_uniform


@dataclass
class Pose:
    p: np.ndarray  # (3,) float32
    q: np.ndarray  # (4,) float32, (x, y, z, w)


IDENTITY = Pose(np.zeros(3, np.float32), np.array([0, 0, 0, 1], np.float32))


@dataclass
class PairWorkload:
    name: str
    K: int
    meshes: list               # Mesh objects
    pairs: np.ndarray          # (npairs, 2) mesh indices (body A, body B)
    poses: list                # one Pose per mesh
    info: dict = field(default_factory=dict)


def single_pair(n=10, K=8, seed=1):
    """Two copies of a procedurally tetrahedralised unit box (n^3 cubes x 5 tets)."""
    m = box_mesh(n, size=1.0, jitter=0.0, seed=seed)
    q = rng_quaternions(seed, 0, 1)[0]
    return PairWorkload(f"box{n}_pair_k{K}", K, [m, m], np.array([[0, 1]], np.uint32),
                        [IDENTITY, Pose(np.array([0.05, 0.05, 0.05], np.float32), q)],
                        dict(n=n, nt=m.nt, nv=m.nv))


def _world_aabb(mesh, pose):
    from . import _quat_rotate_np
    P = _quat_rotate_np(pose.q.astype(np.float64), mesh.points().astype(np.float64)) + pose.p
    return P.min(0), P.max(0)


def aabb_overlap_fraction(lo_a, hi_a, lo_b, hi_b):
    inter = np.clip(np.minimum(hi_a, hi_b) - np.maximum(lo_a, lo_b), 0, None).prod()
    va = (hi_a - lo_a).prod()
    vb = (hi_b - lo_b).prod()
    return float(inter / min(va, vb))


def large_pair(n=40, K=26, seed=7, jitter=0.1, lo=0.3, hi=0.6):
    """Jittered n^3-cube boxes; B gets a seeded rotation and a translation drawn
    until the world-AABB overlap fraction (intersection / smaller box) is in [lo, hi]."""
    ma = box_mesh(n, size=1.0, jitter=jitter, seed=seed)
    mb = box_mesh(n, size=1.0, jitter=jitter, seed=seed + 1)
    q = rng_quaternions(seed, 1, 1)[0]
    lo_a, hi_a = _world_aabb(ma, IDENTITY)
    u = rng_uniform(seed, 2, 3 * 256).reshape(-1, 3)
    frac = None
    for t in u:
        p = ((t * 2.0 - 1.0) * 0.8).astype(np.float32)
        lo_b, hi_b = _world_aabb(mb, Pose(p, q))
        frac = aabb_overlap_fraction(lo_a, hi_a, lo_b, hi_b)
        if lo <= frac <= hi:
            break
    else:
        raise RuntimeError("no translation gave the requested overlap fraction")
    return PairWorkload(f"box{n}_jitter_pair_k{K}", K, [ma, mb], np.array([[0, 1]], np.uint32),
                        [IDENTITY, Pose(p, q)], dict(n=n, nt=ma.nt, nv=ma.nv, aabb_overlap=frac))


def self_contact(n=20, K=26, seed=3, jitter=0.1):
    """A jittered n^3-cube box against an identical copy at identity: every
    tetrahedron overlaps itself and its neighbours, the densest traversal a
    pair of these meshes can make."""
    m = box_mesh(n, size=1.0, jitter=jitter, seed=seed)
    return PairWorkload(f"box{n}_self_contact_k{K}", K, [m, m], np.array([[0, 1]], np.uint32), [IDENTITY, IDENTITY],
                        dict(n=n, nt=m.nt, nv=m.nv))


def batched_pairs(npairs=512, K=8, seed=11, tmin=1e3, tmax=1e5, separated=0.4, jitter=0.1, lo=0.3, hi=0.6):
    """npairs independent body pairs of jittered boxes (edge 0.25); tet counts
    log-uniform in [tmin, tmax]; body A placed uniformly in the unit box; about
    `separated` of the pairs far apart, the others with a seeded rotation and a
    translation drawn (as for large_pair) until the world-AABB overlap fraction
    is in [lo, hi]."""
    u = rng_uniform(seed, 0, 2 * npairs)
    counts = np.exp(np.log(tmin) + u * (np.log(tmax) - np.log(tmin)))
    ncubes = np.maximum(1, np.round((counts / 5.0) ** (1.0 / 3.0))).astype(int)
    cache = {}
    meshes, poses = [], []
    qs = rng_quaternions(seed, 1, 2 * npairs)
    sep = rng_uniform(seed, 2, npairs) < separated
    offs = rng_uniform(seed, 3, 3 * npairs).reshape(npairs, 3)
    fracs = []
    pairs = np.zeros((npairs, 2), np.uint32)
    for i in range(npairs):
        for s in range(2):
            nc = int(ncubes[2 * i + s])
            key = (nc, seed + (2 * i + s) % 7)
            if key not in cache:
                cache[key] = box_mesh(nc, size=0.25, jitter=jitter, seed=key[1])
            meshes.append(cache[key])
        base = (offs[i] * 0.75).astype(np.float32)
        pa = Pose(base, qs[2 * i])
        if sep[i]:
            pb = Pose((base + np.array([2.0, 0.0, 0.0])).astype(np.float32), qs[2 * i + 1])
        else:
            lo_a, hi_a = _world_aabb(meshes[2 * i], pa)
            draws = rng_uniform(seed, 16 + i, 3 * 256).reshape(-1, 3)
            best, pb = None, None
            for t in draws:
                cand = Pose((base + (t * 2.0 - 1.0) * 0.3).astype(np.float32), qs[2 * i + 1])
                lo_b, hi_b = _world_aabb(meshes[2 * i + 1], cand)
                f = aabb_overlap_fraction(lo_a, hi_a, lo_b, hi_b)
                if best is None or abs(f - 0.5 * (lo + hi)) < abs(best - 0.5 * (lo + hi)):
                    best, pb = f, cand
                if lo <= f <= hi:
                    break
            fracs.append(best)
        poses.append(pa)
        poses.append(pb)
        pairs[i] = (2 * i, 2 * i + 1)
    return PairWorkload(f"batched{npairs}_k{K}", K, meshes, pairs, poses,
                        dict(tets=[m.nt for m in meshes], separated=int(sep.sum()),
                             aabb_overlap=(float(np.min(fracs)), float(np.max(fracs))) if fracs else None))


def motion(nbodies, steps, seed=23, dt=1e-3, drift=0.01):
    """Seeded rigid motion: q <- normalize(q * exp(w dt / 2)), w ~ N(0,1) rad/s per
    axis (constant per body), p += drift * (unit direction) per step."""
    q = rng_quaternions(seed, 0, nbodies).astype(np.float64)
    w = rng_normal(seed, 1, 3 * nbodies).reshape(nbodies, 3)
    d = rng_normal(seed, 2, 3 * nbodies).reshape(nbodies, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = (rng_uniform(seed, 3, 3 * nbodies).reshape(nbodies, 3) - 0.5) * 0.1
    out_p = np.empty((steps, nbodies, 3), np.float32)
    out_q = np.empty((steps, nbodies, 4), np.float32)
    for s in range(steps):
        ang = np.linalg.norm(w, axis=1) * dt / 2
        ax = w / np.maximum(np.linalg.norm(w, axis=1, keepdims=True), 1e-30)
        dq = np.concatenate([ax * np.sin(ang)[:, None], np.cos(ang)[:, None]], axis=1)
        q = _qmul(q, dq)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        p = p + drift * d
        out_p[s] = p
        out_q[s] = q
    return out_p, out_q


def _qmul(a, b):
    """Hamilton product of (x, y, z, w) quaternions, row-wise."""
    ax, ay, az, aw = a.T
    bx, by, bz, bw = b.T
    return np.stack([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz], axis=1)


def refit_series(npairs=64, K=18, steps=200, seed=41, tmin=2e4, tmax=8e4, dt=1e-3, drift=0.01):
    """Refit + re-traverse time series: npairs overlapping body pairs (tet
    counts log-uniform in [tmin, tmax], poses of batched_pairs with no
    separated pairs), then `steps` steps of seeded rigid motion per body:
    constant angular velocity w ~ N(0, 1) rad/s per axis, q(s) = q0 * exp(w s dt / 2),
    and a drift of `drift` per step along a seeded unit direction.
    Returns (workload at step 0, P (steps, nbodies, 3), Q (steps, nbodies, 4))."""
    wl = batched_pairs(npairs=npairs, K=K, seed=seed, tmin=tmin, tmax=tmax, separated=0.0)
    nb = len(wl.meshes)
    w = rng_normal(seed, 5, 3 * nb).reshape(nb, 3)
    d = rng_normal(seed, 6, 3 * nb).reshape(nb, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p0 = np.stack([p.p for p in wl.poses]).astype(np.float64)
    q0 = np.stack([p.q for p in wl.poses]).astype(np.float64)
    wn = np.linalg.norm(w, axis=1)
    ax = w / np.maximum(wn[:, None], 1e-30)
    P = np.empty((steps, nb, 3), np.float32)
    Q = np.empty((steps, nb, 4), np.float32)
    for s in range(steps):
        ang = wn * (s * dt) / 2
        dq = np.concatenate([ax * np.sin(ang)[:, None], np.cos(ang)[:, None]], axis=1)
        q = _qmul(q0, dq)
        Q[s] = q / np.linalg.norm(q, axis=1, keepdims=True)
        P[s] = p0 + drift * s * d
    wl.name = f"refit_series{npairs}_k{K}"
    return wl, P, Q
