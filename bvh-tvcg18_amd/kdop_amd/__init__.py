"""Python harness over the MI355X kDOP library (libkdop_amd.so, C ABI).

Mirrors the reference's KDOP call shapes (diku-dk/bvh-tvcg18):

* ``Context.make_tree(mesh)``            ~ kdop::make_tree (kdop_make_tree.h:190-288)
* ``Context.refit(bodies, pos, quat)``   ~ narrow::update_kdop_bvh (narrow_update_kdop_bvh.h:126-176)
* ``Context.tandem_traversal(pairs)``    ~ kdop::tandem_traversal (kdop_tandem_traversal.h:198-226)
* ``box_mesh`` / ``mesh_reorder`` / ``compute_surface_map`` prepare inputs as
  narrow::Tetramesh::set_tetramesh_shape does (narrow_shape_types.h:161-210).

Every compute call goes to the HIP library; a missing library raises.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (CONTACT_DTYPE, KDOP_KEEP_ALL_PAIRS, KDOP_NARROW, KDOP_TRAVERSAL_AUTO,  # noqa: F401
                   KDOP_TRAVERSAL_LEVELS, KDOP_TRAVERSAL_PERSISTENT, Config, Counters, lib, ptr)

SUPPORTED_K = (6, 8, 12, 14, 18, 20, 26)


class KdopError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kdop error {code}: {msg}")
        self.code = code


def _is_gpu_tensor(x):
    """A torch tensor in GPU memory (torch is not imported for host arrays)."""
    return getattr(x, "is_cuda", False) is True


def _gpu_f32(x, width):
    """x as a contiguous float32 GPU tensor of shape (n, width)."""
    import torch
    if not _is_gpu_tensor(x):
        raise ValueError("expected a GPU tensor")
    return x.to(torch.float32).reshape(-1, width).contiguous()


@dataclass
class Mesh:
    """Tetrahedral mesh: SoA vertex coordinates, tets (nt x 4, vertices i j k m),
    surface bits (bit v: face opposite vertex v lies on the surface)."""
    X: np.ndarray
    Y: np.ndarray
    Z: np.ndarray
    tets: np.ndarray
    surface: np.ndarray

    @property
    def nv(self):
        return int(self.X.shape[0])

    @property
    def nt(self):
        return int(self.tets.shape[0])

    def points(self):
        return np.stack([self.X, self.Y, self.Z], axis=1)


# ------------------------------------------------------------ mesh preparation
def box_mesh_raw(n, size=1.0, jitter=0.0, seed=0):
    """Box grid of n^3 cubes x 5 tets in grid order (no reorder, no surface map)."""
    L = lib()
    nv = ctypes.c_uint32()
    nt = ctypes.c_uint32()
    L.kdop_box_mesh_size(n, ctypes.byref(nv), ctypes.byref(nt))
    X = np.empty(nv.value, np.float32)
    Y = np.empty(nv.value, np.float32)
    Z = np.empty(nv.value, np.float32)
    T = np.empty((nt.value, 4), np.uint32)
    rc = L.kdop_make_box_mesh(n, float(size), float(jitter), int(seed), ptr(X), ptr(Y), ptr(Z), ptr(T))
    if rc:
        raise KdopError(rc, "kdop_make_box_mesh failed")
    return X, Y, Z, T


def mesh_reorder(X, Y, Z, tets):
    """Morton reorder of tets, first-touch vertex renumbering (kdop_mesh_reorder.h:172-353)."""
    L = lib()
    nv, nt = X.shape[0], tets.shape[0]
    Xo = np.empty(nv, np.float32)
    Yo = np.empty(nv, np.float32)
    Zo = np.empty(nv, np.float32)
    To = np.empty((nt, 4), np.uint32)
    tets = np.ascontiguousarray(tets, np.uint32)
    rc = L.kdop_mesh_reorder(nv, ptr(X), ptr(Y), ptr(Z), nt, ptr(tets), ptr(Xo), ptr(Yo), ptr(Zo), ptr(To))
    if rc:
        raise KdopError(rc, "kdop_mesh_reorder failed")
    return Xo, Yo, Zo, To


def compute_surface_map(nv, tets):
    """Surface bits per tet (mesh_array_compute_surface_map.h:41-76)."""
    L = lib()
    tets = np.ascontiguousarray(tets, np.uint32)
    out = np.empty(tets.shape[0], np.uint8)
    rc = L.kdop_compute_surface_map(nv, tets.shape[0], ptr(tets), ptr(out))
    if rc:
        raise KdopError(rc, "kdop_compute_surface_map failed")
    return out


def box_mesh(n, size=1.0, jitter=0.0, seed=0, reorder=True):
    """Morton-ordered box-grid mesh with its surface map."""
    X, Y, Z, T = box_mesh_raw(n, size, jitter, seed)
    if reorder:
        X, Y, Z, T = mesh_reorder(X, Y, Z, T)
    S = compute_surface_map(X.shape[0], T)
    return Mesh(X, Y, Z, T, S)


def tree_shape(K, chunk_bytes, nt):
    """Reference chunked-tree shape: dict(node_bytes, n_perfect, L, C, H)."""
    out = np.zeros(5, np.uint64)
    if lib().kdop_tree_shape(K, chunk_bytes, nt, ptr(out)):
        raise KdopError(-1, f"no tree shape for K={K} chunk_bytes={chunk_bytes} nt={nt}")
    return dict(zip(("node_bytes", "n_perfect", "L", "C", "H"), (int(v) for v in out)))


def rng_uniform(seed, stream, n):
    out = np.empty(n, np.float64)
    lib().kdop_rng_uniform(seed, stream, n, ptr(out))
    return out


def rng_normal(seed, stream, n):
    out = np.empty(n, np.float64)
    lib().kdop_rng_normal(seed, stream, n, ptr(out))
    return out


def rng_quaternions(seed, stream, n):
    out = np.empty((n, 4), np.float32)
    lib().kdop_rng_quaternions(seed, stream, n, ptr(out))
    return out


# -------------------------------------------------------------------- context
class Context:
    """One HIP stream + device-resident bodies and trees (kdop_create)."""

    def __init__(self, K=8, chunk_bytes=8000, device=0, traversal=KDOP_TRAVERSAL_AUTO, queue_depth=0, waves_per_cu=0,
                 ramp_target=0, keep=0, seed_budget=0):
        if K not in SUPPORTED_K:
            raise ValueError(f"K={K} not in {SUPPORTED_K}")
        self._lib = lib()
        cfg = Config()
        self._lib.kdop_default_config(ctypes.byref(cfg))
        cfg.K = K
        cfg.chunk_bytes = chunk_bytes
        cfg.device = device
        cfg.traversal = traversal
        cfg.queue_depth = queue_depth
        cfg.waves_per_cu = waves_per_cu
        cfg.ramp_target = ramp_target
        cfg.stack_keep = keep
        cfg.seed_budget = seed_budget
        h = ctypes.c_void_p()
        rc = self._lib.kdop_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc:
            raise KdopError(rc, "kdop_create failed (no HIP device?)")
        self._h = h
        self.K = K
        self.meshes = []
        self._last_npairs = 0
        self.device = device
        self.generation = 0  # bumped by every refit and traversal (the adjoint's inputs)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kdop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise KdopError(rc, self._lib.kdop_last_error(self._h).decode())

    # -- options of the live context (kdop_set_option / kdop_get_option)
    def set_option(self, name, value):
        self._check(self._lib.kdop_set_option(self._h, _abi.OPTIONS[name], int(value)))

    def get_option(self, name):
        v = ctypes.c_int64()
        self._check(self._lib.kdop_get_option(self._h, _abi.OPTIONS[name], ctypes.byref(v)))
        return v.value

    def options(self):
        return {k: self.get_option(k) for k in _abi.OPTIONS}

    def autotune(self, pairs, narrow=True, budget_ms=2000.0):
        """kdop_autotune on this pair list: returns (chosen options, best median ms)."""
        P = np.ascontiguousarray(np.asarray(pairs, np.uint32).reshape(-1, 2))
        chosen = np.zeros(20, np.int64)
        best = ctypes.c_double()
        self._check(self._lib.kdop_autotune(self._h, P.shape[0], ptr(P), KDOP_NARROW if narrow else 0, float(budget_ms),
                                            ptr(chosen), ctypes.byref(best)))
        self._last_npairs = P.shape[0]
        self.generation += 1
        return {k: int(chosen[v - 1]) for k, v in _abi.OPTIONS.items()}, best.value

    # -- kdop::make_tree
    def make_tree(self, mesh, mem_bytes=None):
        """kdop::make_tree(mem_bytes, mesh, ...): mem_bytes None takes the
        context's chunk_bytes (kdop_add_body), else this body's own
        (kdop_add_body_ex; SIZE_MAX = 2**64 - 1 gives one chunk of every tet)."""
        X = np.ascontiguousarray(mesh.X, np.float32)
        Y = np.ascontiguousarray(mesh.Y, np.float32)
        Z = np.ascontiguousarray(mesh.Z, np.float32)
        T = np.ascontiguousarray(mesh.tets, np.uint32)
        S = np.ascontiguousarray(mesh.surface, np.uint8)
        bid = ctypes.c_uint32()
        if mem_bytes is None:
            rc = self._lib.kdop_add_body(self._h, X.shape[0], ptr(X), ptr(Y), ptr(Z), T.shape[0], ptr(T), ptr(S), ctypes.byref(bid))
        else:
            rc = self._lib.kdop_add_body_ex(self._h, int(mem_bytes), X.shape[0], ptr(X), ptr(Y), ptr(Z), T.shape[0], ptr(T), ptr(S),
                                            ctypes.byref(bid))
        self._check(rc)
        self.meshes.append(mesh)
        return bid.value

    def shape(self, body):
        out = np.zeros(5, np.uint64)
        self._check(self._lib.kdop_body_shape(self._h, body, ptr(out)))
        return dict(zip(("L", "C", "H", "nodes", "depth"), (int(v) for v in out)))

    # -- narrow::update_kdop_bvh
    def refit(self, bodies, pos, quat):
        self.generation += 1
        b = np.ascontiguousarray(np.atleast_1d(bodies), np.uint32)
        p = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 3))
        q = np.ascontiguousarray(np.asarray(quat, np.float32).reshape(-1, 4))
        assert p.shape[0] == b.shape[0] == q.shape[0]
        self._check(self._lib.kdop_refit(self._h, b.shape[0], ptr(b), ptr(p), ptr(q)))

    def refit_device(self, bodies, pos_ptr, quat_ptr):
        self.generation += 1
        b = np.ascontiguousarray(np.atleast_1d(bodies), np.uint32)
        self._check(self._lib.kdop_refit_device(self._h, b.shape[0], ptr(b), ctypes.c_void_p(pos_ptr),
                                                ctypes.c_void_p(quat_ptr)))

    # -- kdop::tandem_traversal
    def tandem_traversal(self, pairs, narrow=True, keep_all_pairs=False):
        P = np.ascontiguousarray(np.asarray(pairs, np.uint32).reshape(-1, 2))
        flags = (KDOP_NARROW if narrow else 0) | (KDOP_KEEP_ALL_PAIRS if keep_all_pairs else 0)
        c = Counters()
        self._check(self._lib.kdop_traverse(self._h, P.shape[0], ptr(P), flags, ctypes.byref(c)))
        self._last_npairs = P.shape[0]
        self.generation += 1
        return c.as_dict()

    # -- adjoint of the penetration measure (kdop_narrow_adjoint)
    def narrow_adjoint(self, bodies, dphi=None):
        """Phi per body pair of the last KDOP_NARROW traversal and dL/d(pose)
        of L = sum(dphi * Phi) for `bodies` (frozen choices, FP64 inside):
        (phi (npairs,), grad_pos (n, 3), grad_quat (n, 4)) as float32 torch
        tensors on the context's device.  dphi: torch tensor (npairs,) or None
        (all ones)."""
        import torch
        ids = np.ascontiguousarray(np.asarray(bodies, np.uint32).reshape(-1))
        n, npairs = ids.shape[0], self._last_npairs
        dev = torch.device("cuda", self.device)
        phi = torch.empty(max(npairs, 1), dtype=torch.float32, device=dev)
        gp = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        gq = torch.empty((max(n, 1), 4), dtype=torch.float32, device=dev)
        dp = None
        if dphi is not None:
            dphi = dphi.detach().to(device=dev, dtype=torch.float32).contiguous()
            if dphi.numel() != npairs:
                raise ValueError(f"dphi has {dphi.numel()} entries, the last traversal {npairs} pairs")
            dp = dphi.data_ptr()
        torch.cuda.current_stream(dev).synchronize()  # torch work producing dphi is done
        self._check(self._lib.kdop_narrow_adjoint(self._h, n, ptr(ids), dp, phi.data_ptr(), gp.data_ptr(), gq.data_ptr()))
        return phi[:npairs], gp[:n], gq[:n]

    def pairs(self, which=1):
        """Leaf pairs of the last traversal as an (n, 3) uint32 array (pair, tet_a, tet_b)."""
        n = ctypes.c_uint64()
        self._check(self._lib.kdop_get_pairs(self._h, which, None, 0, 0, ctypes.byref(n)))
        out = np.empty((n.value, 3), np.uint32)
        if n.value:
            self._check(self._lib.kdop_get_pairs(self._h, which, ptr(out), n.value, 0, ctypes.byref(n)))
        return out

    def contacts(self):
        n = ctypes.c_uint64()
        self._check(self._lib.kdop_get_contacts(self._h, None, 0, 0, ctypes.byref(n)))
        out = np.empty(n.value, CONTACT_DTYPE)
        if n.value:
            self._check(self._lib.kdop_get_contacts(self._h, ptr(out), n.value, 0, ctypes.byref(n)))
        return out

    def sort_contacts(self, order="sequential"):
        """Sort the last traversal's contacts on the device (kdop_sort_contacts):
        "sequential" = (pair, tet_a, tet_b, ordinal), the sequential reference's
        report order; "opencl" = (pair, p, n, depth), the OpenCL path's qsort order."""
        code = {"sequential": _abi.KDOP_ORDER_SEQUENTIAL, "opencl": _abi.KDOP_ORDER_OPENCL}[order]
        self._check(self._lib.kdop_sort_contacts(self._h, code))

    def reduce_contacts(self):
        """PROX contact reduction of the last traversal's contacts on the device
        (kdop_reduce_contacts): sorted by (body i, body j, p.y, p.x, p.z, n.y, n.x,
        n.z, depth), contacts within the working precision of the last one kept
        dropped.  Returns the count left; contacts() then returns them."""
        n = ctypes.c_uint64()
        self._check(self._lib.kdop_reduce_contacts(self._h, ctypes.byref(n)))
        return n.value

    def reduce_prox_contacts(self, contacts):
        """The same reduction over an array of PROX_CONTACT_DTYPE records
        (kdop_reduce_prox_contacts); returns the reduced records."""
        c = np.ascontiguousarray(contacts, _abi.PROX_CONTACT_DTYPE)
        out = np.zeros(c.shape[0], _abi.PROX_CONTACT_DTYPE)
        n = ctypes.c_uint64()
        self._check(self._lib.kdop_reduce_prox_contacts(self._h, c.shape[0], ptr(c) if c.shape[0] else None,
                                                        ptr(out) if c.shape[0] else None, 0, ctypes.byref(n)))
        return out[:n.value]

    def mesh_reorder(self, X, Y, Z, tets):
        """Morton reorder of a mesh on the device (kdop_mesh_reorder_gpu): (X, Y, Z, tets) reordered."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        Z = np.ascontiguousarray(Z, np.float32)
        T = np.ascontiguousarray(np.asarray(tets, np.uint32).reshape(-1, 4))
        nv, nt = X.shape[0], T.shape[0]
        Xo, Yo, Zo = (np.empty(nv, np.float32) for _ in range(3))
        To = np.empty((nt, 4), np.uint32)
        self._check(self._lib.kdop_mesh_reorder_gpu(self._h, nv, ptr(X), ptr(Y), ptr(Z), nt, ptr(T), ptr(Xo), ptr(Yo), ptr(Zo),
                                                    ptr(To)))
        return Xo, Yo, Zo, To

    def raycast(self, body, origins, directions):
        """kdop_raycast: (hit points (n, 3), lengths (n,), hit flags (n,) bool) of n rays at `body`.
        With GPU tensors for origins and directions the call reads and writes
        device memory in place and returns GPU tensors (hit flags uint8)."""
        if _is_gpu_tensor(origins) or _is_gpu_tensor(directions):
            return self._raycast_device(body, origins, directions)
        o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        n = o.shape[0]
        if d.shape[0] != n:
            raise ValueError("origins and directions differ in length")
        hits = np.zeros((n, 3), np.float32)
        lengths = np.zeros(n, np.float32)
        flags = np.zeros(n, np.uint8)
        self._check(self._lib.kdop_raycast(self._h, body, n, ptr(o), ptr(d), ptr(hits), ptr(lengths), ptr(flags), 0))
        return hits, lengths, flags.astype(bool)

    def _raycast_device(self, body, origins, directions):
        import torch
        o = _gpu_f32(origins, 3)
        d = _gpu_f32(directions, 3)
        n = o.shape[0]
        if d.shape[0] != n or d.device != o.device:
            raise ValueError("origins and directions differ in length or device")
        hits = torch.empty((n, 3), dtype=torch.float32, device=o.device)
        lengths = torch.empty(n, dtype=torch.float32, device=o.device)
        flags = torch.empty(n, dtype=torch.uint8, device=o.device)
        if n:
            torch.cuda.current_stream(o.device).synchronize()  # inputs written on torch's stream are complete
            self._check(self._lib.kdop_raycast(self._h, body, n, o.data_ptr(), d.data_ptr(), hits.data_ptr(), lengths.data_ptr(),
                                               flags.data_ptr(), 1))
        return hits, lengths, flags

    def single_traversal(self, body, spheres, flip=False):
        """kdop_single_traversal: contacts of spheres (n, 4: x, y, z, radius) with `body`,
        grouped by sphere, each sphere's in the reference's report order.
        With a GPU tensor of spheres the call reads it in place and returns
        the records as a GPU uint8 tensor (count, record size) in device memory."""
        if _is_gpu_tensor(spheres):
            return self._single_traversal_device(body, spheres, flip)
        sp = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 4))
        n = ctypes.c_uint64()
        # room for 64 contacts per sphere (or the last call's count) first, so
        # that one call usually suffices; np.empty touches no memory
        cap = max(64 * sp.shape[0], getattr(self, "_single_last", 0), 1)
        out = np.empty(cap, _abi.SPHERE_CONTACT_DTYPE)
        rc = self._lib.kdop_single_traversal(self._h, body, sp.shape[0], ptr(sp), int(flip), ptr(out), cap, ctypes.byref(n), 0)
        if rc == _abi.KDOP_ERR_CAPACITY:
            out = np.empty(n.value, _abi.SPHERE_CONTACT_DTYPE)
            rc = self._lib.kdop_single_traversal(self._h, body, sp.shape[0], ptr(sp), int(flip), ptr(out), n.value, ctypes.byref(n), 0)
        self._check(rc)
        self._single_last = n.value
        return out[:n.value]

    def _single_traversal_device(self, body, spheres, flip):
        import torch
        sp = _gpu_f32(spheres, 4)
        rec = _abi.SPHERE_CONTACT_DTYPE.itemsize
        n = ctypes.c_uint64()
        cap = max(64 * sp.shape[0], getattr(self, "_single_last", 0), 1)
        out = torch.empty((cap, rec), dtype=torch.uint8, device=sp.device)
        if sp.shape[0]:
            torch.cuda.current_stream(sp.device).synchronize()
        rc = self._lib.kdop_single_traversal(self._h, body, sp.shape[0], sp.data_ptr(), int(flip), out.data_ptr(), cap,
                                             ctypes.byref(n), 1)
        if rc == _abi.KDOP_ERR_CAPACITY:
            out = torch.empty((n.value, rec), dtype=torch.uint8, device=sp.device)
            rc = self._lib.kdop_single_traversal(self._h, body, sp.shape[0], sp.data_ptr(), int(flip), out.data_ptr(), n.value,
                                                 ctypes.byref(n), 1)
        self._check(rc)
        self._single_last = n.value
        return out[:n.value]

    def world(self, body):
        nv = self.meshes[body].nv
        X = np.empty(nv, np.float32)
        Y = np.empty(nv, np.float32)
        Z = np.empty(nv, np.float32)
        self._check(self._lib.kdop_get_world(self._h, body, ptr(X), ptr(Y), ptr(Z)))
        return X, Y, Z

    def export_tree(self, body):
        n = ctypes.c_uint64()
        self._check(self._lib.kdop_export_tree(self._h, body, 0, None, None, None, None, ctypes.byref(n)))
        vol = np.empty((n.value, self.K), np.float32)
        rng = np.empty((n.value, 2), np.uint32)
        kind = np.empty(n.value, np.uint8)
        root = np.empty(self.K, np.float32)
        self._check(self._lib.kdop_export_tree(self._h, body, n.value, ptr(vol), ptr(rng), ptr(kind), ptr(root),
                                               ctypes.byref(n)))
        return dict(vol=vol, range=rng, kind=kind, root=root)

    def traversal_stats(self):
        out = np.zeros(20, np.uint64)
        self._check(self._lib.kdop_traversal_stats(self._h, ptr(out)))
        keys = ("ramp_levels", "frontier", "rounds", "iters", "lanes", "max_iters", "grabs", "waves",
                "persist_tests", "persist_launches", "queue_depth", "waves_per_cu", "handoffs", "item_lanes",
                "block_waves", "fallback_levels", "records_loaded", "distinct_records")
        return dict(zip(keys, (int(v) for v in out)))

    def gather_roof(self, window=0, nloads=1 << 23, seed=1):
        """kdop_gather_roof: GB/s of random record loads over the first `window`
        node slots of this context's pool (0: all of it)."""
        g = ctypes.c_double()
        self._check(self._lib.kdop_gather_roof(self._h, int(window), int(nloads), int(seed), ctypes.byref(g)))
        return g.value

    def last_timing(self):
        out = np.zeros(12, np.float64)
        self._check(self._lib.kdop_last_timing(self._h, ptr(out)))
        return dict(traversal_ms=out[0], narrow_ms=out[1], total_ms=out[2], per_launch_ms=out[3],
                    round_ms=[float(v) for v in out[4:]])


def _quat_rotate_np(q, P):
    """Rotate points P (n x 3) by quaternion q = (x, y, z, w) in float64 (for
    workload set-up only: pose sampling, AABB estimates)."""
    u = np.asarray(q[:3], np.float64)
    s = float(q[3])
    P = np.asarray(P, np.float64)
    uv = np.cross(u, P)
    return (s * s - u @ u) * P + 2.0 * np.outer(P @ u, u) + 2.0 * s * uv
