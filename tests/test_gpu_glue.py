"""The C++ call shapes of include/kdop_amd.hpp (make_tree, a pool refit,
a pair-list tandem traversal with one ContactsCallback per pair) driven by a
C++ program, tests/glue/glue_test.cpp: every callback must receive exactly the
oracle's contacts for its pair, in the oracle's report order, bit for bit.
The program is built by `python bvh-tvcg18_amd/build.py --oracle`."""
import os
import subprocess

import numpy as np
import pytest

from kdop_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "glue", "glue_test")


def _binary():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run `python bvh-tvcg18_amd/build.py --oracle`")
    return BIN


def test_cpp_call_shapes_match_oracle():
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.count("match") == 10 and "MISMATCH" not in r.stdout


# 16 overlapping pairs of 2-6k-tet jittered boxes (K=18) under 50 steps of
# seeded rigid motion, slow enough that every pair stays in contact
SERIES = dict(npairs=16, K=18, steps=50, seed=41, tmin=2000, tmax=6000, drift=0.0005)


def write_series(path, wl, P, Q):
    """The series file glue_test --series reads (its run_series describes the layout)."""
    with open(path, "wb") as f:
        f.write(np.array([wl.K, len(wl.meshes), len(wl.pairs), P.shape[0]], np.int32).tobytes())
        for m in wl.meshes:
            f.write(np.array([m.X.shape[0], m.tets.shape[0]], np.uint32).tobytes())
            for a in (m.X, m.Y, m.Z):
                f.write(np.ascontiguousarray(a, np.float32).tobytes())
            f.write(np.ascontiguousarray(m.tets, np.uint32).tobytes())
            f.write(np.ascontiguousarray(m.surface, np.uint8).tobytes())
        f.write(np.ascontiguousarray(wl.pairs, np.uint32).tobytes())
        f.write(np.ascontiguousarray(P, np.float32).tobytes())
        f.write(np.ascontiguousarray(Q, np.float32).tobytes())


def test_cpp_prox_loop_over_a_motion_series(tmp_path):
    """PROX's collision-detection loop (prox_collision_detection.h:124-345)
    through kdop_amd.hpp, step after step: pool refit, pair-list traversal
    with a callback per pair, contact reduction.  Every step, every pair's
    callback contacts equal the oracle's (count, order, bits), and the
    reduction keeps what the oracle's std::sort + std::unique keeps."""
    wl, P, Q = W.refit_series(**SERIES)
    path = str(tmp_path / "series.bin")
    write_series(path, wl, P, Q)
    r = subprocess.run([_binary(), "--series", path], capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-4000:], r.stderr[-2000:])
    assert r.stdout.count(" match\n") == SERIES["steps"] and "MISMATCH" not in r.stdout
