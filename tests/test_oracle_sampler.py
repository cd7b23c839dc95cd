"""Pins the oracle's B-spline sampler, FFT, prefilter and noise generator.

Checks are independent of the oracle's own formulas: the cardinal cubic
B-spline closed form, SciPy's map_coordinates/spline_filter (grid-wrap),
numpy's FFT and a direct O(n^2) DFT, and the reference's noise_golden fixture
(reference tests/test_filtered_noise.cpp:68-324).
"""
import json
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def beta3(x):
    ax = np.abs(x)
    return np.where(ax < 1, 2.0 / 3 - ax ** 2 + ax ** 3 / 2,
                    np.where(ax < 2, (2 - ax) ** 3 / 6, 0.0))


def dbeta3(x):
    ax = np.abs(x)
    s = np.sign(x)
    return s * np.where(ax < 1, -2 * ax + 1.5 * ax ** 2,
                        np.where(ax < 2, -0.5 * (2 - ax) ** 2, 0.0))


def test_bspline_weights_closed_form():
    import ctypes as C
    lib = O.load()
    for t in np.linspace(0, 1, 41, endpoint=False):
        w = (C.c_double * 4)()
        dw = (C.c_double * 4)()
        lib.orc_bspline_weights(t, w, dw)
        offs = np.arange(4) - 1.0
        np.testing.assert_allclose(list(w), beta3(t - offs), atol=1e-15)
        np.testing.assert_allclose(list(dw), dbeta3(t - offs), atol=1e-14)
        assert abs(sum(w) - 1) < 1e-15 and abs(sum(dw)) < 1e-14


@pytest.mark.parametrize("nx,ny,Lx,Ly", [(16, 12, 1.6, 1.0), (7, 9, 2.5, 1.3), (32, 32, 32.0, 32.0)])
def test_sampler_matches_scipy_map_coordinates(nx, ny, Lx, Ly):
    rng = np.random.default_rng(nx * 100 + ny)
    coef = rng.standard_normal((nx, ny))
    nz = O.OracleNoise.table(coef, Lx, Ly)
    xq = rng.uniform(-3 * Lx, 4 * Lx, 300)
    yq = rng.uniform(-3 * Ly, 4 * Ly, 300)
    v, dvx, dvy = O.sample_points(nz, xq, yq)
    gx = np.mod(xq / (Lx / nx), nx)
    gy = np.mod(yq / (Ly / ny), ny)
    ref = ndimage.map_coordinates(coef, [gx, gy], order=3, mode="grid-wrap", prefilter=False)
    np.testing.assert_allclose(v, ref, atol=1e-12)
    # derivatives: brute force sum of the tensor basis with the closed-form beta3'
    ix = np.floor(gx).astype(int)
    iy = np.floor(gy).astype(int)
    dref_x = np.zeros_like(v)
    dref_y = np.zeros_like(v)
    for a in range(-1, 3):
        for b in range(-1, 3):
            c = coef[(ix + a) % nx, (iy + b) % ny]
            dref_x += c * dbeta3(gx - ix - a) * beta3(gy - iy - b)
            dref_y += c * beta3(gx - ix - a) * dbeta3(gy - iy - b)
    np.testing.assert_allclose(dvx, dref_x / (Lx / nx), atol=1e-10)
    np.testing.assert_allclose(dvy, dref_y / (Ly / ny), atol=1e-10)


def test_value_and_dy_samplers_equal_full_sample():
    rng = np.random.default_rng(5)
    coef = rng.standard_normal((20, 18))
    nz = O.OracleNoise.table(coef, 2.0, 1.0)
    lib = O.load()
    for x, y in rng.uniform(-5, 5, (50, 2)):
        v, dx, dy = O.sample_points(nz, [x], [y])
        assert lib.orc_sample_value(nz.ref, x, y) == v[0]
        assert lib.orc_sample_dy(nz.ref, x, y) == dy[0]


def test_periodic_wrap_across_many_cells():
    # reference tests/test_filtered_noise.cpp:238-281
    field, coef = O.filtered_noise(48, 40, 2.5, 1.3, 1.0, 0.25, 0.25, 17)
    nz = O.OracleNoise.table(coef, 2.5, 1.3)
    rng = np.random.default_rng(123)
    xs, ys = rng.uniform(0, 2.5, 60), rng.uniform(0, 1.3, 60)
    v0 = O.sample_points(nz, xs, ys)
    for sx, sy in [(1, 0), (0, 1), (-1, 0), (0, -1), (3, -2), (-5, 7), (-13, -11)]:
        v = O.sample_points(nz, xs + sx * 2.5, ys + sy * 1.3)
        for a, b in zip(v, v0):
            assert np.max(np.abs(a - b)) < 1e-9


@pytest.mark.parametrize("n", [1, 2, 3, 7, 48, 64, 100, 256, 1024])
def test_fft_matches_direct_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    got = O.fft(x, -1)
    np.testing.assert_allclose(got, dft, atol=1e-9 * max(1, n))
    np.testing.assert_allclose(got, np.fft.fft(x), atol=1e-12 * max(1, n))
    back = O.fft(got, +1) / n
    assert np.max(np.abs(back - x)) < 1e-12  # tests/test_fft_util.cpp:7-11


def test_fft2_backward_normalised():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))
    np.testing.assert_allclose(O.fft2(a, -1), np.fft.fft2(a), atol=1e-12)
    np.testing.assert_allclose(O.fft2(a, +1), np.fft.ifft2(a), atol=1e-15)


@pytest.mark.parametrize("nx,ny", [(8, 6), (48, 32), (64, 48), (15, 10)])
def test_prefilter_matches_scipy_and_solves_circulant(nx, ny):
    rng = np.random.default_rng(nx + ny)
    field = rng.standard_normal((nx, ny))
    coef = O.prefilter(field)
    np.testing.assert_allclose(coef, ndimage.spline_filter(field, order=3, mode="grid-wrap"),
                               atol=1e-12)
    # (1/6, 2/3, 1/6) circulant in both axes applied to coef reproduces field
    k = lambda c, ax: (np.roll(c, 1, ax) + 4 * c + np.roll(c, -1, ax)) / 6
    np.testing.assert_allclose(k(k(coef, 0), 1), field, atol=1e-12)


def test_noise_golden_fixture():
    # reference tests/test_filtered_noise.cpp:286-324
    g = json.load(open(os.path.join(GOLD, "noise_golden.json")))
    nx, ny = int(g["nx"]), int(g["ny"])
    field = np.array(g["field"]["data"]).reshape(nx, ny)
    nz = O.OracleNoise.table(O.prefilter(field), g["Lx"], g["Ly"])
    v, dvx, dvy = O.sample_points(nz, np.array(g["x"]["data"]), np.array(g["y"]["data"]))
    assert np.max(np.abs(v - g["v"]["data"])) < 1e-7
    assert np.max(np.abs(dvx - g["dvdx"]["data"])) < 1e-4
    assert np.max(np.abs(dvy - g["dvdy"]["data"])) < 1e-4


def test_reproduces_values_at_nodes():
    # reference tests/test_filtered_noise.cpp:118-143
    field, coef = O.filtered_noise(64, 48, 3.2, 1.0, 1.5, 0.15, 0.15, 3)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    ii, jj = np.meshgrid(np.arange(0, 64, 5), np.arange(0, 48, 5), indexing="ij")
    v, _, _ = O.sample_points(nz, ii.ravel() * (3.2 / 64), jj.ravel() * (1.0 / 48))
    assert np.max(np.abs(v - field[ii.ravel(), jj.ravel()])) < 1e-9


def test_analytic_derivatives_match_fd():
    # reference tests/test_filtered_noise.cpp:147-181
    _, coef = O.filtered_noise(64, 48, 3.2, 1.0, 1.5, 0.15, 0.15, 5)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    rng = np.random.default_rng(99)
    xs, ys = rng.uniform(0, 3.2, 100), rng.uniform(0, 1.0, 100)
    ex, ey = 1e-5 * 3.2 / 64, 1e-5 / 48
    _, dvx, dvy = O.sample_points(nz, xs, ys)
    fdx = (O.sample_points(nz, xs + ex, ys)[0] - O.sample_points(nz, xs - ex, ys)[0]) / (2 * ex)
    fdy = (O.sample_points(nz, xs, ys + ey)[0] - O.sample_points(nz, xs, ys - ey)[0]) / (2 * ey)
    assert np.max(np.abs(fdx - dvx)) < 1e-5 and np.max(np.abs(fdy - dvy)) < 1e-5


# ---- noise generator: mt19937_64 restated independently in Python --------

class MT19937_64:
    """std::mt19937_64 (Matsumoto & Nishimura 2000; C++ [rand.predef])."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & (2 ** 64 - 1)
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & (2 ** 64 - 1)
        self.idx = 312

    def __call__(self):
        M = 2 ** 64 - 1
        if self.idx >= 312:
            for i in range(312):
                x = (self.mt[i] & 0xFFFFFFFF80000000) | (self.mt[(i + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[i] = self.mt[(i + 156) % 312] ^ xa
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000 & M
        y ^= (y << 37) & 0xFFF7EEE000000000 & M
        y ^= y >> 43
        return y


def test_mt19937_64_restatement_known_answer():
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042  # C++ [rand.predef] requirement


def _python_filtered_noise(nx, ny, Lx, Ly, amp, xi_x, xi_y, seed):
    """src/filtered_noise.cpp:25-82 restated with numpy's FFT."""
    dx, dy = Lx / nx, Ly / ny
    two_pi = 2 * np.arccos(-1.0)
    g = MT19937_64(seed)
    A = np.zeros((nx, ny), complex)
    ff = lambda k, n: (k if k < (n + 1) // 2 else k - n) / n
    for i in range(nx):
        fx = ff(i, nx) / dx
        for j in range(ny):
            fy = ff(j, ny) / dy
            u = float(g()) / 2.0 ** 64  # generate_canonical<double, 53> on a 64-bit engine
            if u >= 1.0:
                u = np.nextafter(1.0, 0.0)
            ph = two_pi * u
            if (fx * xi_x) ** 2 + (fy * xi_y) ** 2 <= 1.0:
                A[i, j] = np.cos(ph) + 1j * np.sin(ph)
    B = np.fft.ifft2(A).real
    return B * (amp / np.std(B))


def test_filtered_noise_generator_matches_python_restatement():
    field, _ = O.filtered_noise(24, 20, 1.2, 1.0, 1.7, 0.15, 0.1, 7)
    ref = _python_filtered_noise(24, 20, 1.2, 1.0, 1.7, 0.15, 0.1, 7)
    np.testing.assert_allclose(field, ref, atol=1e-12)


def test_filtered_noise_amplitude_and_band_limit():
    # reference tests/test_filtered_noise.cpp:68-114
    nx, ny, Lx, Ly, amp, xx, xy = 128, 96, 4.0, 1.0, 2.0, 0.2, 0.1
    field, _ = O.filtered_noise(nx, ny, Lx, Ly, amp, xx, xy, 42)
    assert abs(field.std() - amp) < 1e-10
    H = np.abs(np.fft.fft2(field))
    fx = np.fft.fftfreq(nx) / (Lx / nx)
    fy = np.fft.fftfreq(ny) / (Ly / ny)
    Fx, Fy = np.meshgrid(fx, fy, indexing="ij")
    masked = (Fx * xx) ** 2 + (Fy * xy) ** 2 > 1
    assert H[masked].max() < 1e-8 * H.max()


def _mid_fold(coef):
    """The x-sum of the bicubic at x-fraction 1/2 per (ix, iy), FP64: the
    cardinal cubic B-spline's values at offsets -3/2, -1/2, 1/2, 3/2 times
    coef(ix - 1 .. ix + 2)."""
    w = beta3(np.array([1.5, 0.5, -0.5, -1.5]))
    return sum(w[a] * np.roll(coef, 1 - a, axis=0) for a in range(4))


@pytest.mark.parametrize("nx,ny,dx,dy", [(48, 40, 1.0, 1.0), (64, 32, 0.05, 0.25)])
def test_oracle_midpoint_table_equals_full_sampler_at_half_cells(nx, ny, dx, dy):
    """The oracle's MIDTABLE sampler (T = the x-contracted table) gives the
    full bicubic's V and dV/dy at every point with x-fraction 1/2, including
    the cell of the wrap edge ((n - 1) / 2), for any y; the weights w(+-1/2),
    w(+-3/2) are (23, 1)/48. A column off by one or a transposed table fails."""
    rng = np.random.default_rng(nx)
    coef = rng.standard_normal((nx, ny))
    full = O.OracleNoise.table(coef, nx * dx, ny * dy)
    mid = O.OracleNoise.mid_table(_mid_fold(coef), nx * dx, ny * dy)
    np.testing.assert_allclose(beta3(np.array([1.5, 0.5])) * 48, [1, 23], rtol=0, atol=1e-13)
    i = np.concatenate([np.arange(nx - 1), [nx - 1]])
    ip = (i + 1) % nx
    x = (i + ip) * 0.5 * dx
    y = rng.uniform(-2 * ny * dy, 3 * ny * dy, x.size)
    v, _, dv = O.sample_points(full, x, y)
    vm, dxm, dvm = O.sample_points(mid, x, y)
    scale = np.max(np.abs(v))
    np.testing.assert_allclose(vm, v, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(dvm, dv, rtol=0, atol=1e-13 * np.max(np.abs(dv)))
    assert np.all(np.isnan(dxm))
    # a planted off-by-one column shift is caught
    bad = O.OracleNoise.mid_table(np.roll(_mid_fold(coef), 1, axis=0), nx * dx, ny * dy)
    vb, _, _ = O.sample_points(bad, x, y)
    assert np.max(np.abs(vb - v)) > 1e-3 * scale


@pytest.mark.parametrize("model", [0, 1])
def test_gradient_edges_are_the_sampler_and_slope(model):
    """The oracle's per-edge workspace values are dV/dy of the (separately
    pinned) sampler at each edge midpoint, including the wrap edge's at
    (n - 1) dx / 2, and the arclength slope s / sqrt(1 + s^2); the gradient
    assembled from them is orc_gradient's."""
    n, ny = 40, 32
    field, coef = O.filtered_noise(n, ny, 2.0, 1.0, 0.5, 0.2, 0.2, 9)
    on = O.OracleNoise.table(coef, 2.0, 1.0)
    h = 0.5 + 0.05 * np.random.default_rng(2).standard_normal(n)
    p = O.params(2.0, 1.1, 0.05, model)
    g, e, s = O.gradient_edges(h, p, on)
    dx = 2.0 / n
    i = np.arange(n)
    ip = (i + 1) % n
    lib = O.load()
    e_ref = np.array([lib.orc_sample_dy(on.ref, (a + b) * 0.5 * dx, (h[b] + h[a]) * 0.5)
                      for a, b in zip(i, ip)])
    np.testing.assert_array_equal(e, e_ref)
    if model == 1:
        sl = (h[ip] - h) / dx
        np.testing.assert_allclose(s, sl / np.sqrt(1 + sl * sl), rtol=1e-15, atol=0)
    else:
        assert s is None
    np.testing.assert_array_equal(g, O.gradient(h, p, on))
