"""The C++ front-end (include/RAJA) with the host policies: seq_exec and
omp_parallel_for_exec forall, reducers, expt params, scans, sorts and
WorkGroup, checked against closed forms (no GPU needed)."""
import os
import subprocess

from tests.cpp import build as cppbuild


def test_frontend_host_policies():
    exe = cppbuild.build_host()
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
