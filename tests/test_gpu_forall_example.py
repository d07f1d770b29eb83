"""The forall + reducers workload (raja_amd/examples/daxpy_dot_minloc.cpp) on
the GPU: daxpy bit-exact (no contraction), dot within 2*ceil(log2 N)*eps*sum|x*y|
of a compensated host sum for both old-style reducers and expt::Reduce
parameters (and the two agree bit for bit), MinLoc the first index of the
minimum."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "raja_amd", "bin", "daxpy_dot_minloc")


@pytest.mark.parametrize("log2n", [10, 17, 24])
def test_daxpy_dot_minloc(log2n):
    if not os.path.exists(EXE):
        pytest.fail("raja_amd/bin/daxpy_dot_minloc is not built")
    print("daxpy_dot_minloc", log2n, flush=True)
    r = subprocess.run([EXE, str(log2n), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["pass"] and out["daxpy_bit_exact"]
    assert out["dot_abs_err_vs_kn"] <= out["dot_bound"]
