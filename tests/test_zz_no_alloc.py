"""spc2_noise_noalloc, the entry point that allocates no device memory: its
normals equal spc2_noise's bit for bit, on two calls in a row of which the
second stages fewer normals than the first (the staging array keeps the first
call's tail), and sizes beyond its fixed array are refused."""
import numpy as np
import pytest

import spc2_langevin_integration_amd as spc2

pytestmark = pytest.mark.gpu


def params(n, seed):
    return spc2.params("chimera", "fp32k", n=n, L=float(np.sqrt(n / 4000.0)), rho=0.1, seed=seed)


def test_noise_noalloc_equals_noise_on_consecutive_calls():
    spc2.load()
    calls = [(1048576, "srk", 9, 12345), (5000, "se", 10, 77)]   # 2^21 normals, then 5000
    with spc2.Sim(params(calls[0][0], calls[0][2])) as big, spc2.Sim(params(calls[1][0], calls[1][2])) as small:
        ref = [big.noise(calls[0][1], calls[0][3]), small.noise(calls[1][1], calls[1][3])]
        got = [big.noise_noalloc(calls[0][1], calls[0][3]), small.noise_noalloc(calls[1][1], calls[1][3])]
        again = big.noise_noalloc(calls[0][1], calls[0][3] + 1)
        ref_again = big.noise(calls[0][1], calls[0][3] + 1)
    for (n, method, _, _), a, b in zip(calls, ref, got):
        assert b.size == n * (2 if method == "srk" else 1)
        assert np.array_equal(a, b)
    assert np.array_equal(again, ref_again)


def test_noise_noalloc_refuses_sizes_beyond_its_array():
    spc2.load()
    with spc2.Sim(params(1048577, 3)) as s:
        with pytest.raises(spc2.Spc2Error, match="at most 2097152 normals"):
            s.noise_noalloc("srk", 0)
        assert s.noise_noalloc("se", 0).size == 1048577
