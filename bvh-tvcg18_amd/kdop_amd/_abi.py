"""ctypes binding of libkdop_amd.so (include/kdop_amd.h, include/kdop_meshgen.h)."""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "lib", "libkdop_amd.so")

KDOP_OK = 0
KDOP_ERR_ARG = -1
KDOP_ERR_HIP = -2
KDOP_ERR_NOMEM = -3
KDOP_ERR_STATE = -4
KDOP_ERR_CAPACITY = -5
KDOP_ERR_INTERNAL = -6

KDOP_NOALLOC_MAX_CONTACTS = 262144  # include/kdop_amd.h

KDOP_TRAVERSAL_AUTO = 0
KDOP_TRAVERSAL_LEVELS = 1
KDOP_TRAVERSAL_PERSISTENT = 2

# kdop_set_option / kdop_get_option (include/kdop_amd.h)
OPTIONS = {"traversal": 1, "queue_depth": 2, "waves_per_cu": 3, "ramp_target": 4, "stack_keep": 5,
           "seed_budget": 6, "kernel_timing": 7, "bfs_max": 8, "async_refit": 9,
           "narrow_path": 10, "graphs": 11, "single_items": 12,
           "block_waves": 13, "item_lanes": 14, "max_rounds": 15, "count_records": 16, "quad_items": 17, "round_iters": 18, "node_slab": 19}
KDOP_NARROW = 1
KDOP_ORDER_SEQUENTIAL = 0
KDOP_ORDER_OPENCL = 1
KDOP_KEEP_ALL_PAIRS = 2


class Config(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("chunk_bytes", ctypes.c_uint64),
        ("traversal", ctypes.c_int32),
        ("queue_depth", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("ramp_target", ctypes.c_int32),
        ("stack_keep", ctypes.c_int32),
        ("seed_budget", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class Counters(ctypes.Structure):
    _fields_ = [
        ("tests", ctypes.c_uint64),
        ("reference_tests", ctypes.c_uint64),
        ("descent_tests", ctypes.c_uint64),
        ("branch_overlaps", ctypes.c_uint64),
        ("leaf_pairs", ctypes.c_uint64),
        ("surface_pairs", ctypes.c_uint64),
        ("contacts", ctypes.c_uint64),
        ("narrow_pairs", ctypes.c_uint64),
        ("branch_pairs", ctypes.c_uint64),
        ("overflow", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


# numpy view of kdop_sphere_contact (36 bytes)
SPHERE_CONTACT_DTYPE = np.dtype([
    ("sphere", np.uint32), ("tet", np.uint32), ("p", np.float32, (3,)), ("n", np.float32, (3,)), ("depth", np.float32),
])
# numpy view of kdop_prox_contact (36 bytes)
PROX_CONTACT_DTYPE = np.dtype([
    ("body_i", np.uint32), ("body_j", np.uint32), ("p", np.float32, (3,)), ("n", np.float32, (3,)), ("depth", np.float32),
])
# numpy view of kdop_contact (44 bytes)
CONTACT_DTYPE = np.dtype([
    ("pair", np.uint32), ("tet_a", np.uint32), ("tet_b", np.uint32), ("ordinal", np.uint32),
    ("p", np.float32, (3,)), ("n", np.float32, (3,)), ("depth", np.float32),
])

_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32

# (name, restype, argtypes) for every entry point of include/*.h
SIGNATURES = [
    ("kdop_default_config", None, [_p]),
    ("kdop_create", ctypes.c_int, [_p, _p]),
    ("kdop_destroy", None, [_p]),
    ("kdop_last_error", ctypes.c_char_p, [_p]),
    ("kdop_add_body", ctypes.c_int, [_p, _u32, _p, _p, _p, _u32, _p, _p, _p]),
    ("kdop_add_body_ex", ctypes.c_int, [_p, _u64, _u32, _p, _p, _p, _u32, _p, _p, _p]),
    ("kdop_body_shape", ctypes.c_int, [_p, _u32, _p]),
    ("kdop_refit", ctypes.c_int, [_p, _u32, _p, _p, _p]),
    ("kdop_refit_device", ctypes.c_int, [_p, _u32, _p, _p, _p]),
    ("kdop_traverse", ctypes.c_int, [_p, _u32, _p, _u32, _p]),
    ("kdop_get_pairs", ctypes.c_int, [_p, ctypes.c_int, _p, _u64, ctypes.c_int, _p]),
    ("kdop_get_contacts", ctypes.c_int, [_p, _p, _u64, ctypes.c_int, _p]),
    ("kdop_sort_contacts", ctypes.c_int, [_p, ctypes.c_int]),
    ("kdop_reduce_contacts", ctypes.c_int, [_p, _p]),
    ("kdop_reduce_prox_contacts", ctypes.c_int, [_p, _u64, _p, _p, ctypes.c_int, _p]),
    ("kdop_reduce_prox_contacts_noalloc", ctypes.c_int, [_p, _u64, _p, _p, _p]),
    ("kdop_gather_roof", ctypes.c_int, [_p, _u64, _u32, _u32, _p]),
    ("kdop_raycast", ctypes.c_int, [_p, _u32, _u32, _p, _p, _p, _p, _p, ctypes.c_int]),
    ("kdop_single_traversal", ctypes.c_int, [_p, _u32, _u32, _p, ctypes.c_int, _p, _u64, _p, ctypes.c_int]),
    ("kdop_mesh_reorder_gpu", ctypes.c_int, [_p, _u32, _p, _p, _p, _u32, _p, _p, _p, _p, _p]),
    ("kdop_get_world", ctypes.c_int, [_p, _u32, _p, _p, _p]),
    ("kdop_export_tree", ctypes.c_int, [_p, _u32, _u64, _p, _p, _p, _p, _p]),
    ("kdop_narrow_adjoint", ctypes.c_int, [_p, _u32, _p, _p, _p, _p, _p]),
    ("kdop_last_timing", ctypes.c_int, [_p, _p]),
    ("kdop_traversal_stats", ctypes.c_int, [_p, _p]),
    ("kdop_set_option", ctypes.c_int, [_p, ctypes.c_int, ctypes.c_int64]),
    ("kdop_autotune", ctypes.c_int, [_p, _u32, _p, _u32, ctypes.c_double, _p, _p]),
    ("kdop_get_option", ctypes.c_int, [_p, ctypes.c_int, _p]),
    ("kdop_stream", _p, [_p]),
    ("kdop_box_mesh_size", None, [_u32, _p, _p]),
    ("kdop_make_box_mesh", ctypes.c_int, [_u32, ctypes.c_float, ctypes.c_double, _u64, _p, _p, _p, _p]),
    ("kdop_mesh_reorder", ctypes.c_int, [_u32, _p, _p, _p, _u32, _p, _p, _p, _p, _p]),
    ("kdop_compute_surface_map", ctypes.c_int, [_u32, _u32, _p, _p]),
    ("kdop_tree_shape", ctypes.c_int, [_i32, _u64, _u64, _p]),
    ("kdop_rng_uniform", None, [_u64, _u64, _u64, _p]),
    ("kdop_rng_normal", None, [_u64, _u64, _u64, _p]),
    ("kdop_rng_quaternions", None, [_u64, _u64, _u64, _p]),
]

_lib = None


def lib():
    """Load libkdop_amd.so (built in-tree by bvh-tvcg18_amd/build.py).

    PyTorch-ROCm wheels carry their own HIP runtime under the same soname
    (libamdhip64.so.7).  torch is imported first when it is installed, so the
    library binds to that already loaded runtime and one process never holds
    two HIP runtimes.  Set KDOP_NO_TORCH=1 to skip the import.

    Raises RuntimeError when the library is missing: there is no fallback path.
    """
    global _lib
    if _lib is not None:
        return _lib
    if os.environ.get("KDOP_NO_TORCH", "0") != "1":
        try:
            import torch  # noqa: F401  (binds the shared HIP runtime first)
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python bvh-tvcg18_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a):
    """ctypes pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C contiguous"
    return ctypes.c_void_p(a.ctypes.data)
