"""CPU-side checks of the C ABI library: it loads, and exports every entry
point the public headers declare (no device calls)."""
import ctypes
import os
import re

import kdop_amd as K
from kdop_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(kdop_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    L = K.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in _abi.SIGNATURES}
    assert set(names) == bound, (set(names) ^ bound)


def test_default_config_and_error_on_missing_device():
    L = K.lib()
    cfg = _abi.Config()
    L.kdop_default_config(ctypes.byref(cfg))
    assert cfg.K == 8 and cfg.chunk_bytes == 8000 and cfg.device == 0


def test_tree_shape_matches_quoted_workload_shapes():
    # chunks C and levels H quoted for the reference formula (kdop_make_tree.h:216-229)
    assert K.tree_shape(8, 8000, 5000) == dict(node_bytes=56, n_perfect=63, L=32, C=157, H=2)
    s = K.tree_shape(26, 8000, 320000)
    assert (s["L"], s["C"], s["H"]) == (8, 40000, 6)
    s = K.tree_shape(6, 8000, 320000)
    assert (s["L"], s["C"], s["H"]) == (32, 10000, 3)
    for k in (14, 18):
        s = K.tree_shape(k, 8000, 320000)
        assert (s["L"], s["C"], s["H"]) == (16, 20000, 4)
    s = K.tree_shape(14, 8000, 50000)
    assert (s["L"], s["C"], s["H"]) == (16, 3125, 3)
    # chunks wider than 32 leaves (a per-body mem_bytes): K=8 nodes are 56 B, so
    # 16000 / 2 / 56 = 142 -> 128 slots -> 64 leaves; 32000 -> 256 slots -> 128
    # leaves; SIZE_MAX: the smallest power of two >= 2 M slots, one chunk
    assert K.tree_shape(8, 16000, 5000) == dict(node_bytes=56, n_perfect=127, L=64, C=79, H=2)
    assert K.tree_shape(8, 32000, 5000) == dict(node_bytes=56, n_perfect=255, L=128, C=40, H=1)
    assert K.tree_shape(8, 2 ** 64 - 1, 5000) == dict(node_bytes=56, n_perfect=16383, L=8192, C=1, H=1)


def test_cpp_header_compiles():
    """include/kdop_amd.hpp compiles on its own with the host compiler."""
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with tempfile.NamedTemporaryFile("w", suffix=".cpp", delete=False) as f:
        f.write('#include "kdop_amd.hpp"\nint main() { kdop_amd::Tree t; return (int)(t.body & 0u); }\n')
        src = f.name
    try:
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(root, "include"), src],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    finally:
        os.unlink(src)


def test_record_layouts_match_numpy_views():
    """Each public record of include/kdop_amd.h has the size and field offsets of
    its numpy view (static_asserts compiled by the host compiler)."""
    import subprocess
    import tempfile
    from kdop_amd import _abi
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stddef.h>', '#include "kdop_amd.h"']
    for ctype, dt in (("kdop_contact", _abi.CONTACT_DTYPE), ("kdop_sphere_contact", _abi.SPHERE_CONTACT_DTYPE),
                      ("kdop_prox_contact", _abi.PROX_CONTACT_DTYPE)):
        lines.append(f'static_assert(sizeof({ctype}) == {dt.itemsize}, "{ctype} size");')
        for name in dt.names:
            lines.append(f'static_assert(offsetof({ctype}, {name}) == {dt.fields[name][1]}, "{ctype}.{name}");')
    with tempfile.NamedTemporaryFile("w", suffix=".cpp", delete=False) as f:
        f.write("\n".join(lines) + "\nint main() { return 0; }\n")
        src = f.name
    try:
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(root, "include"), src],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    finally:
        os.unlink(src)
