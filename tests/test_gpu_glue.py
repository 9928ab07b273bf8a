"""The C++ call shapes of include/kdop_amd.hpp (make_tree, a pool refit, a
pair-list tandem traversal with one ContactsCallback per pair) driven by a
C++ program, tests/glue/glue_test.cpp: every callback must receive exactly the
oracle's contacts for its pair, in the oracle's report order, bit for bit.
The program is built by `python bvh-tvcg18_amd/build.py --oracle`."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "glue", "glue_test")


def test_cpp_call_shapes_match_oracle():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run `python bvh-tvcg18_amd/build.py --oracle`")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.count("match") == 10 and "MISMATCH" not in r.stdout
