"""The WorkGroup halo exchange (raja_amd/examples/halo_exchange.cpp) on the
GPU: its neighbour lists equal the oracle's restatement of the reference's
create_pack_lists / create_unpack_lists, and every variant (per-loop forall,
ordered WorkGroup, unordered tile WorkGroup) leaves every cell of every
variable as a sequential pack + unpack does."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "raja_amd", "bin", "halo_exchange")


def run_halo(*args, timeout=120):
    if not os.path.exists(EXE):
        pytest.fail("raja_amd/bin/halo_exchange is not built")
    print("halo_exchange", *args, flush=True)
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def list_hash(lists):
    i = np.arange(1, lists.size + 1, dtype=np.uint64)
    return int(np.sum(lists.astype(np.uint64) * i, dtype=np.uint64))


@pytest.mark.parametrize("dims,h,nvars", [((20, 24, 28), 2, 3), ((9, 7, 5), 1, 2),
                                          ((12, 12, 12), 3, 1)])
def test_halo_lists_and_variants(orc, dims, h, nvars):
    out = run_halo(*dims, h, nvars, 2)
    lens_p, pl = orc.halo_lists(list(dims), h, 0)
    lens_u, ul = orc.halo_lists(list(dims), h, 1)
    assert out["list_total"] == int(np.sum(lens_p)) == int(np.sum(lens_u))
    assert out["pack_hash"] == list_hash(pl)
    assert out["unpack_hash"] == list_hash(ul)
    for v in ("forall", "ordered", "unordered"):
        assert out[v]["pass"], v


def test_halo_config_size(golden):
    out = run_halo(256, 256, 256, 2, 10, 2, timeout=300)
    assert out["list_total"] == sum(golden["halo"]["pack"]["lengths"]) == 798784
    assert out["loops_per_phase"] == 260
    for v in ("forall", "ordered", "unordered"):
        assert out[v]["pass"], v
