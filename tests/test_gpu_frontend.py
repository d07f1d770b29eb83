"""The C++ front-end with the cdna4 policies on the GPU: forall with every
old-style reducer (init offsets, accumulation, reset, first-index MinLoc on
tied data), expt::Reduce / ReduceLoc parameters, daxpy bit-exact against the
host, ListSegment and strided ranges, async resources with an event
dependency, scans and sorts against seq_exec, and WorkGroup runs (unordered
tile runner, ordered, reverse ordered; indirect and direct dispatch; halo
pack/unpack round trip).  The program is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_frontend_gpu")


def test_frontend_cdna4_policies():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/bin/test_frontend_gpu is not built; run __graft_entry__.build()")
    env = dict(os.environ, OMP_NUM_THREADS="8")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
