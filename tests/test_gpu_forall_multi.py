"""One forall with ReduceSum/Min/Max/MinLoc over FP64 N(0,1)
(raja_amd/examples/forall_multi_reducers.cpp) on the GPU: min, max and the
first index of the minimum exact, the sum within 2*ceil(log2 N)*eps*sum|x| of
a compensated host sum, old-style reducers, expt::Reduce parameters and the
loop whose body reads x through a nontemporal load bit for bit identical."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "raja_amd", "bin", "forall_multi_reducers")


@pytest.mark.parametrize("log2n", [4, 13, 20, 26, 28])
def test_forall_multi_reducers(log2n):
    if not os.path.exists(EXE):
        pytest.fail("raja_amd/bin/forall_multi_reducers is not built")
    print("forall_multi_reducers", log2n, flush=True)
    r = subprocess.run([EXE, str(log2n), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["pass"] and out["styles_bit_identical"]
    assert out["reducers_nontemporal_body"]["same_bits"]
    assert out["sum_abs_err_vs_kn"] <= out["sum_bound"]
    assert out["minloc"][0] == out["min"]
