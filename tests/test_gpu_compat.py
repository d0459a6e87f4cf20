"""GPU: a caller written against the reference's utils API (tests/compat),
linked to the compatible shim libraries, gets the FFT of its data."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "compat", "_build")


@pytest.mark.parametrize("args,dim", [(["-n", "256", "-b", "64"], 1), (["-n", "131072", "-b", "2"], 1),
                                      (["-n", "4096", "-b", "3"], 1), (["-x", "256", "-y", "512", "-b", "2"], 2),
                                      (["-x", "64", "-y", "32", "-z", "128", "-b", "2"], 3),
                                      (["-x", "256", "-y", "256", "-b", "2"], 3)])
def test_reference_api_caller(torch_cuda, args, dim):
    exe = os.path.join(BIN, f"accuracy_{dim}d")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    print(r)
    # generate_data() values are ~1e-4 (partly FP16-subnormal); error vs the FP64
    # FFT of the quantised input stays at the plain-FP16 level
    assert r["rel_l2"] < 2e-3
    assert r["status"] == 0
