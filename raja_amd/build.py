"""Build libraja_cdna4.so (gfx950) in-tree with hipcc.

Every HIP source under raja_amd/csrc is compiled for gfx950 only and linked
into raja_amd/libraja_cdna4.so.  Packed-FP32 instruction selection is
disabled and floating-point contraction is off: a*x+y is a multiply then an
add, the same two roundings a CPU computes.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libraja_cdna4.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    objs = []
    failed = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(_compile, srcs):
            # hipcc's host pass does not know the packed-fp32 feature; drop that line.
            log = "\n".join(l for l in log.splitlines()
                            if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
            if rc != 0:
                failed.append((src, log))
            elif verbose and log.strip():
                print(log)
            objs.append(obj)
    if failed:
        for src, log in failed:
            sys.stderr.write("=== %s\n%s\n" % (src, log))
        raise RuntimeError("HIP compilation failed: %s" % ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


EXAMPLES = os.path.join(HERE, "examples")
BIN = os.path.join(HERE, "bin")


def build_examples():
    """Example programs of the C++ front-end (raja_amd/examples/*.cpp) into
    raja_amd/bin/, linked to libraja_cdna4.so through an $ORIGIN rpath."""
    os.makedirs(BIN, exist_ok=True)
    outs = []
    for name in sorted(os.listdir(EXAMPLES)):
        if not name.endswith(".cpp"):
            continue
        out = os.path.join(BIN, name[:-4])
        cmd = [HIPCC, "-std=c++17", "-O3", "-ffp-contract=off", "-Wall"] + DEVICE + [
            "-I" + os.path.join(ROOT, "include"), os.path.join(EXAMPLES, name),
            "-L" + HERE, "-lraja_cdna4", "-Wl,-rpath,$ORIGIN/..", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s build failed:\n%s%s" % (name, r.stdout, r.stderr))
        outs.append(out)
    return outs


PLUGINS = os.path.join(HERE, "plugins")


def build_plugins():
    """Tool plugins (raja_amd/plugins/*.cpp, host code) as RAJA_PLUGINS libraries."""
    outs = []
    for name in sorted(os.listdir(PLUGINS)):
        if not name.endswith(".cpp"):
            continue
        out = os.path.join(PLUGINS, "lib" + name[:-4] + ".so")
        cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
               "-I/opt/rocm/include", os.path.join(PLUGINS, name), "-L/opt/rocm/lib",
               "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s build failed:\n%s%s" % (name, r.stdout, r.stderr))
        outs.append(out)
    return outs


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print("\n".join(build_examples()))
    print("\n".join(build_plugins()))
