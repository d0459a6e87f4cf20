"""GPU: the autotuner picks a plan structure and a kernel per pass, persists
them under the library's build ID, and a later plan of the same shape starts
from the cached choice; results stay correct.  A cache written by another build
is ignored."""
import os
import subprocess
import sys

import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def kernel_names(p):
    return [l.split()[1] for l in p.describe().splitlines()[1:]]


def structure(p):
    return p.describe().splitlines()[0].split("struct=")[1]


@pytest.mark.parametrize("dims,batch", [([65536], 16), ([1 << 18], 8), ([1024, 1024], 8)])
def test_tune_roundtrip_cache(torch_cuda, tmp_path, monkeypatch, dims, batch):
    torch = torch_cuda
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("MCFFT_TUNE_CACHE", str(cache))
    n = int(np.prod(dims))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)
    p = mc.Plan(dims, batch, "fp16", "fp16")
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(n * batch, "fp16")
    p.tune(d_in, d_out, mc.FORWARD, reps=3)
    assert cache.exists()
    lines = cache.read_text().splitlines()
    assert lines[0] == "mcfft-tune-cache 2 " + mc.build_id() and len(lines) >= 2
    print(dims, "tuned:", structure(p), kernel_names(p))
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    y = p.result(d_out, 2)
    assert oracle.rel_l2(y, oracle.fft(xq[:2 * n], tuple(dims))) <= 1.5e-3
    # a new plan of the same shape reads the cache: same structure and kernels
    q = mc.Plan(dims, batch, "fp16", "fp16")
    assert structure(q) == structure(p)
    assert kernel_names(q) == kernel_names(p)


def test_stale_build_cache_ignored(torch_cuda, tmp_path):
    """An entry that would change a 2D plan's structure, in a file of another
    build, leaves a new process's plan at its defaults; the same entry under this
    build's ID is taken."""
    sig = "gfx950|2|1024x1024|8|0|0|3"
    entry = sig + "=R=0;L2=0;K=mcfft_pass_2_4_2_4_0_0,mcfft_pass_2_4_32_16_0_2\n"
    prog = ("import sys; sys.path.insert(0, %r); import mcfft_amd as mc; "
            "print(mc.Plan([1024, 1024], 8, 'fp16', 'fp16').describe().splitlines()[0])" % ROOT)
    out = {}
    for tag, bid in (("stale", "0000000000000000"), ("current", mc.build_id())):
        cache = tmp_path / (tag + ".txt")
        cache.write_text("mcfft-tune-cache 2 " + bid + "\n" + entry)
        env = dict(os.environ, MCFFT_TUNE_CACHE=str(cache))
        r = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out[tag] = r.stdout.strip().splitlines()[-1]
    assert "struct=R=4;" in out["stale"], out     # the default for 1024 x 1024
    assert "struct=R=0;" in out["current"], out
