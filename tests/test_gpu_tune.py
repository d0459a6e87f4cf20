"""GPU: the autotuner picks a kernel per pass, persists it, and a later plan of
the same shape starts from the cached choice; results stay correct."""
import os

import numpy as np
import pytest

import mcfft_amd as mc
import oracle

pytestmark = pytest.mark.gpu


def test_tune_roundtrip_cache(torch_cuda, tmp_path, monkeypatch):
    torch = torch_cuda
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("MCFFT_TUNE_CACHE", str(cache))
    n, b = 65536, 16
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n * b) + 1j * rng.standard_normal(n * b)
    p = mc.Plan([n], b, "fp16", "fp16")
    d_in = mc.to_device(x, "fp16")
    d_out = mc.empty(n * b, "fp16")
    p.tune(d_in, d_out, mc.FORWARD, reps=3)
    assert cache.exists()
    lines = cache.read_text().splitlines()
    assert lines[0] == "mcfft-tune-cache 1" and len(lines) >= 2
    p.exec(d_in, d_out, mc.FORWARD)
    torch.cuda.synchronize()
    xq = mc.decode(mc.encode(x, "fp16"), "fp16")
    y = p.result(d_out)
    assert oracle.rel_l2(y, oracle.fft(xq, (n,))) <= 1.5e-3
    # a new plan of the same shape reads the cache: same pass kernels
    q = mc.Plan([n], b, "fp16", "fp16")
    assert [l.split()[1] for l in q.describe().splitlines()[1:]] == \
           [l.split()[1] for l in p.describe().splitlines()[1:]]
