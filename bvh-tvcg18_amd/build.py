"""Build the MI355X kDOP library (libkdop_amd.so) in-tree with hipcc for gfx950.

    python bvh-tvcg18_amd/build.py            # library only
    python bvh-tvcg18_amd/build.py --oracle   # also the CPU oracle (tests only)

Floating-point flags matter for parity: -ffp-contract=off keeps every a*b+c
as two rounded operations (the reference's g++ build has no FMA contraction),
-fhip-fp32-correctly-rounded-divide-sqrt makes device sqrt and division
correctly rounded like the host's, and packed FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libkdop_amd.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KDOP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off", "-fno-fast-math",
    "-I", INCLUDE, "-I", CSRC,
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_library():
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    hip_src = os.path.join(CSRC, "kdop_hip.hip")
    hip_obj = os.path.join(LIBDIR, "kdop_hip.o")
    run([HIPCC] + COMMON + DEVICE + ["-c", hip_src, "-o", hip_obj])
    objs.append(hip_obj)
    for src in ("meshgen.cpp",):
        obj = os.path.join(LIBDIR, src.replace(".cpp", ".o"))
        run([HIPCC] + COMMON + ["-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_oracle():
    run(["make", "-C", os.path.join(ROOT, "oracle")])


def build_glue_test():
    """tests/glue/glue_test: a C++ caller of include/kdop_amd.hpp, checked against the oracle."""
    src = os.path.join(ROOT, "tests", "glue", "glue_test.cpp")
    out = os.path.join(ROOT, "tests", "glue", "glue_test")
    run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, src, "-o", out,
         "-L", LIBDIR, "-lkdop_amd", "-L", os.path.join(ROOT, "oracle"), "-loracle",
         f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath,{os.path.join(ROOT, 'oracle')}", "-Wl,-rpath,$ORIGIN/../../bvh-tvcg18_amd/lib",
         "-Wl,-rpath,$ORIGIN/../../oracle", "-lpthread"])


def main(argv):
    build_library()
    if "--oracle" in argv:
        build_oracle()
        build_glue_test()


if __name__ == "__main__":
    main(sys.argv[1:])
