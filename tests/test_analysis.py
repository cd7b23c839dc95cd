"""Roughness analysis (qew_amd.analysis), on the CPU.

The reference's analysis module is not part of its checkout, so these tests
pin the functions by what the mathematics fixes: the autocorrelation against
the brute-force double sum, the PSD against a direct DFT and Parseval's
identity, and the exponent fit on profiles whose exponent is known exactly
(power-law spectra with random phases) or statistically (random walks)."""
import numpy as np
import pytest

from qew_amd import analysis as A


def _brute_acf(h, r):
    n = h.shape[-1]
    x = np.arange(n)
    return 0.5 * np.mean((h[..., (x + r) % n] - h[..., x]) ** 2)


@pytest.mark.parametrize("n", [2, 3, 37, 64])
def test_acf_matches_double_sum(n):
    rng = np.random.default_rng(n)
    h = rng.standard_normal((3, n)) * 2.0 + 5.0
    L = 7.5
    ell, acf = A.height_difference_acf(h, L)
    assert len(ell) == n // 2
    np.testing.assert_allclose(ell, np.arange(1, n // 2 + 1) * L / n, rtol=0, atol=1e-15)
    ref = np.array([_brute_acf(h, r) for r in range(1, n // 2 + 1)])
    np.testing.assert_allclose(acf, ref, rtol=1e-12, atol=1e-13)
    # one profile alone
    ell1, acf1 = A.height_difference_acf(h[1], L)
    np.testing.assert_allclose(acf1, [_brute_acf(h[1], r) for r in range(1, n // 2 + 1)],
                               rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("n", [3, 16, 37])
def test_psd_matches_direct_dft_and_parseval(n):
    rng = np.random.default_rng(100 + n)
    h = rng.standard_normal(n) + 3.0
    L = 2.5
    q, C = A.psd(h, L)
    k = np.arange(1, n // 2 + 1)
    np.testing.assert_allclose(q, 2 * np.pi * k / L, rtol=1e-15)
    x = np.arange(n) * L / n
    direct = np.array([abs(np.sum(h * np.exp(-1j * qq * x))) ** 2 * L / n ** 2 for qq in q])
    np.testing.assert_allclose(C, direct, rtol=1e-12)
    # Parseval: (1/L) sum over k != 0 of C = var(h); k and -k both count,
    # the Nyquist term (even n) once
    weights = np.full(len(C), 2.0)
    if n % 2 == 0:
        weights[-1] = 1.0
    assert abs(np.sum(weights * C) / L - np.var(h)) < 1e-12 * np.var(h)
    assert abs(A.width2(h) - np.var(h)) < 1e-14


def _power_law_profile(n, H, rng, L=1.0):
    """Real periodic profile with |h_k| = k^-(1 + 2H)/2 and random phases."""
    k = np.arange(n // 2 + 1, dtype=np.float64)
    amp = np.zeros_like(k)
    amp[1:] = k[1:] ** (-(1 + 2 * H) / 2)
    spec = amp * np.exp(2j * np.pi * rng.random(len(k)))
    spec[-1] = spec[-1].real if n % 2 == 0 else spec[-1]
    return np.fft.irfft(spec, n=n)


@pytest.mark.parametrize("H", [0.3, 0.5, 0.8])
def test_exponent_of_exact_power_law(H):
    rng = np.random.default_rng(int(H * 10))
    n, L = 4096, 4096.0
    h = _power_law_profile(n, H, rng)
    # the PSD is an exact power law: the fit returns H to rounding
    assert abs(A.roughness_exponent(h, L, 4.0, 1024.0, method="psd") - H) < 1e-9
    # the ACF of a stack of such profiles follows ell^(2H) at small ell
    hs = np.stack([_power_law_profile(n, H, rng) for _ in range(16)])
    assert abs(A.roughness_exponent(hs, L, 2.0, 64.0) - H) < 0.06


def test_exponent_of_random_walks():
    """A random walk (closed into a bridge by the periodic ACF) has
    A(ell) = sigma^2 ell (1 - ell / n) / 2 in expectation: zeta = 1/2."""
    rng = np.random.default_rng(5)
    n = 8192
    steps = rng.standard_normal((64, n))
    steps -= steps.mean(axis=1, keepdims=True)  # periodic bridge
    h = np.cumsum(steps, axis=1)
    ell, acf = A.height_difference_acf(h, float(n))
    expect = 0.5 * ell * (1 - ell / n)
    sel = ell <= 256
    assert np.max(np.abs(acf[sel] / expect[sel] - 1)) < 0.05
    assert abs(A.roughness_exponent(h, float(n), 1.0, 128.0) - 0.5) < 0.02


def test_errors():
    with pytest.raises(ValueError):
        A.height_difference_acf(np.zeros(1), 1.0)
    with pytest.raises(ValueError):
        A.roughness_exponent(np.arange(16.0), 16.0, 1.0, 2.0)
    with pytest.raises(ValueError):
        A.roughness_exponent(np.arange(16.0), 16.0, 1.0, 8.0, method="other")
