"""GPU roughness analysis (qew_roughness through the C ABI) against the host
implementation in qew_amd.analysis, whose conventions the CPU tests pin by
closed forms (tests/test_analysis.py), and against closed forms directly: a
sinusoid of wavenumber k0 has a single PSD line and A(r) = (a^2 / 2)(1 -
cos(2 pi k0 r / n)); a constant profile has zero width, ACF and PSD."""
import numpy as np
import pytest

from gpu_util import dev, host, torch_cuda
from qew_amd import analysis

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


@pytest.mark.parametrize("m,n", [(1, 2), (3, 64), (5, 100), (2, 1026), (4, 4096), (1, 16384)])
def test_roughness_matches_host_analysis(q, m, n):
    rng = np.random.default_rng(n + m)
    L = 0.37 * n
    # random walks (rough) with an offset, as relaxed profiles look
    h = 3.0 + np.cumsum(rng.standard_normal((m, n)), axis=1)
    ell, a, qq, c, w2 = q.roughness(dev(h), L)
    ell_ref, a_ref = analysis.height_difference_acf(h, L)
    q_ref, c_ref = analysis.psd(h, L)
    np.testing.assert_allclose(host(ell), ell_ref, rtol=1e-15)
    np.testing.assert_allclose(host(qq), q_ref, rtol=1e-15)
    # FFT rounding differs between the two paths: compare on the scale of the data
    assert np.max(np.abs(host(a) - a_ref)) <= 1e-12 * np.max(np.abs(a_ref))
    assert np.max(np.abs(host(c) - c_ref)) <= 1e-12 * np.max(np.abs(c_ref))
    np.testing.assert_allclose(host(w2), analysis.width2(h), rtol=1e-12)


def test_roughness_closed_forms(q):
    n, k0, amp = 96, 5, 1.7
    L = 12.0
    x = np.arange(n)
    h = amp * np.sin(2 * np.pi * k0 * x / n) + 0.25
    ell, a, qq, c, w2 = q.roughness(dev(h), L)
    r = np.arange(1, n // 2 + 1)
    np.testing.assert_allclose(host(a), amp ** 2 / 2 * (1 - np.cos(2 * np.pi * k0 * r / n)),
                               atol=1e-13)
    expect = np.zeros(n // 2)
    # |sum_x a sin(2 pi k0 x / n) e^{-2 pi i k0 x / n}|^2 = (a n / 2)^2
    expect[k0 - 1] = (amp * n / 2) ** 2 * L / n ** 2
    np.testing.assert_allclose(host(c), expect, atol=1e-12)
    np.testing.assert_allclose(host(w2), [amp ** 2 / 2], rtol=1e-13)
    # a flat profile
    ell, a, qq, c, w2 = q.roughness(dev(np.full((2, 64), 7.5)), 64.0)
    assert np.all(host(a) == 0) and np.all(host(c) == 0) and np.all(host(w2) == 0)


def test_roughness_outputs_optional_and_bad_input(q):
    torch = torch_cuda()
    h = dev(np.random.default_rng(3).standard_normal((2, 128)))
    ell, a, qq, c, w2 = q.roughness(h, 128.0, acf=False, psd=False)
    assert a is None and c is None
    np.testing.assert_allclose(host(w2), analysis.width2(host(h)), rtol=1e-12)
    with pytest.raises(ValueError):
        q.roughness(dev(np.zeros((2, 1))), 1.0)
    with pytest.raises(ValueError):
        q.roughness(h, -1.0)
