"""Build the C++ front-end test programs in-tree (tests/cpp/bin/).

test_frontend_host: g++ -fopenmp, host policies only (runs without a GPU).
test_frontend_gpu:  hipcc for gfx950 with the cdna4 policies, linked to
                    raja_amd/libraja_cdna4.so through an $ORIGIN rpath.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
SRC = os.path.join(HERE, "test_frontend.cpp")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build_host():
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "test_frontend_host")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-I" + INC, SRC,
                    "-o", out], check=True)
    return out


def build_gpu():
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "test_frontend_gpu")
    cmd = [HIPCC, "-std=c++17", "-O3", "-ffp-contract=off", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fopenmp",
           "-I" + INC, SRC, "-L" + os.path.join(ROOT, "raja_amd"), "-lraja_cdna4",
           "-Wl,-rpath,$ORIGIN/../../../raja_amd", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("test_frontend_gpu build failed:\n" + r.stdout + r.stderr)
    return out


def build_plugin():
    """tests/cpp/bin/libcounter_plugin.so: a RAJA_PLUGINS test plugin (host code)."""
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "libcounter_plugin.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I" + INC,
                    os.path.join(HERE, "counter_plugin.cpp"), "-o", out], check=True)
    return out


if __name__ == "__main__":
    print(build_host())
    print(build_gpu())
    print(build_plugin())
