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


def test_runtime_plugin_loaded_from_raja_plugins():
    """A plugin library named by RAJA_PLUGINS is dlopened and sees every
    forall/launch (src/RuntimePluginLoader.cpp:94-127 in the reference)."""
    exe = cppbuild.build_host()
    lib = cppbuild.build_plugin()
    for path in (lib, os.path.dirname(lib)):  # a library, or a directory of them
        env = dict(os.environ, OMP_NUM_THREADS="4", RAJA_PLUGINS=path)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        line = [l for l in r.stdout.splitlines() if l.startswith("[counter plugin]")]
        assert len(line) == 1, r.stdout[-2000:]
        fields = dict(kv.split("=", 1) for kv in line[0].split()[2:])
        assert int(fields["pre"]) == int(fields["post"]) == int(fields["capture"]) > 10
        assert int(fields["device"]) == 0  # host build: every kernel is a host loop
        assert "plugin-named-loop;" in fields["names"]
