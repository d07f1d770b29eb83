"""The cdna4 customization points have the reference's shapes: compiling
tests/cpp/test_boundary.cpp with hipcc for gfx950 instantiates
param_init/param_combine/param_resolve through expt::ParamMultiplexer,
forall_impl with the parameter pack by value, and detail::WorkRunner's
enqueue/run/clear, and checks their signatures with static_asserts.
Runs on the CPU (hipcc cross-compiles; nothing is executed)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_boundary_shapes_compile(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "test_boundary.cpp")
    out = str(tmp_path / "test_boundary.o")
    r = subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950",
                        "-I" + os.path.join(ROOT, "include"), "-c", src, "-o", out],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.getsize(out) > 0
