"""The multi-tau correlator of the Flag_correlation mode (host code in
libcvf.so; reference gpu_water3D.cu:2983-3128) against brute-force
autocorrelations of the same series.  Level 0 holds lags 0..15 of the raw
series; level k >= 1 holds lags 8..15 of the series averaged over aligned
blocks of 2^k samples; the highest level reached is not reported."""
import numpy as np
import pytest

from cvfbulkwater_amd import Correlator, LIB_PATH

P, M = 16, 2

pytestmark = pytest.mark.skipif(not LIB_PATH.exists(), reason="libcvf.so not built")


def brute(x):
    x = np.asarray(x, float)
    mu2 = x.mean() ** 2
    t, f = [], []
    for j in range(P):
        if j < len(x):
            t.append(j)
            f.append(np.mean(x[j:] * x[: len(x) - j]) - mu2)
    # level k exists once 2^k samples arrived; the top level is left out
    kmax = int(np.floor(np.log2(len(x))))
    for k in range(1, kmax):
        b = x[: len(x) // 2 ** k * 2 ** k].reshape(-1, 2 ** k).mean(axis=1)
        for i in range(P // M, P):
            if i < len(b):
                t.append(i * 2 ** k)
                f.append(np.mean(b[i:] * b[: len(b) - i]) - mu2)
    f = np.array(f)
    return np.array(t, float), f / f[0]


@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (16, 2), (17, 3), (1000, 4), (4096, 5), (12345, 6)])
def test_correlator_matches_brute_force(n, seed):
    rng = np.random.default_rng(seed)
    # an AR(1) series in the range of the largest-sigma fraction (1/6 .. 1)
    x = np.empty(n)
    x[0] = 0.5
    for i in range(1, n):
        x[i] = 0.5 + 0.9 * (x[i - 1] - 0.5) + 0.02 * rng.standard_normal()
    c = Correlator()
    for v in x:
        c.add(v)
    t, f = c.result()
    if n == 1:
        # a single sample: f[0] = x^2 - x^2 = 0, normalised 0/0
        assert list(t) == [0.0] and np.isnan(f[0])
        return
    tb, fb = brute(x)
    assert np.array_equal(t, tb)
    np.testing.assert_allclose(f, fb, rtol=1e-9, atol=1e-9)


def test_correlator_known_series():
    """Alternating +-1 around 0.5: lag j correlation is (-1)^j * 0.25 at
    level 0; every block average at level >= 1 equals the mean, so those
    lags are exactly 0."""
    x = 0.5 + 0.5 * (-1.0) ** np.arange(256)
    c = Correlator()
    for v in x:
        c.add(v)
    t, f = c.result()
    assert list(t[:16]) == list(range(16))
    np.testing.assert_allclose(f[:16], (-1.0) ** np.arange(16), atol=1e-12)
    np.testing.assert_allclose(f[16:], 0.0, atol=1e-12)
    # level 4 holds 16 block means, the last with lag-15 pairs; levels 5..7
    # (8, 4, 2 means) have none of the lags 8..15, level 8 is the top
    assert t[-1] == 15 * 2 ** 4


def test_correlator_empty():
    t, f = Correlator().result()
    assert len(t) == 0 and len(f) == 0
