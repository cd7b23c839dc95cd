"""GPU parity of the per-point entry points the reference's own tests call
(through the C ABI): DeviceNoise::sample_value / sample_dy
(src/filtered_noise.h:138-176) and rk_forward_zero
(src/rosso_krauth.h:88-124), bit for bit against the oracle on the same
tables; the reference's sampler and site-velocity tests
(tests/test_filtered_noise.cpp:118-281, tests/test_rosso_krauth.cpp:94-141)
restated through the C++ facade live in tests/cpp/facade_selftest.cpp."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


def test_sample_value_and_dy_match_oracle_bitwise(q):
    nx, ny, Lx, Ly = 48, 40, 2.5, 1.3
    field, coef = O.filtered_noise(nx, ny, Lx, Ly, 1.0, 0.25, 0.25, 17)
    rng = np.random.default_rng(5)
    # points inside the box, on nodes, and many periods out (both signs)
    x = np.concatenate([rng.uniform(0, Lx, 200), np.arange(nx) * (Lx / nx),
                        rng.uniform(-13 * Lx, 11 * Lx, 100)])
    y = np.concatenate([rng.uniform(0, Ly, 200), np.arange(nx) % ny * (Ly / ny),
                        rng.uniform(-7 * Ly, 9 * Ly, 100)])
    for c in (coef, coef.astype(np.float32).astype(np.float64)):
        dtype = q.F64 if c is coef else q.F32
        gn = q.FilteredNoise.from_coefficients(coef, Lx / nx, Ly / ny, dtype=dtype)
        on = O.OracleNoise.table(c, Lx, Ly)
        v = host(gn.sample_value(dev(x), dev(y)))
        d = host(gn.sample_dy(dev(x), dev(y)))
        lib = O.load()
        v_ref = np.array([lib.orc_sample_value(on.ref, a, b) for a, b in zip(x, y)])
        d_ref = np.array([lib.orc_sample_dy(on.ref, a, b) for a, b in zip(x, y)])
        np.testing.assert_array_equal(v, v_ref)
        np.testing.assert_array_equal(d, d_ref)
        # the full sampler's value and dV/dy are the same numbers
        vf, _, df = gn.sample(dev(x), dev(y))
        np.testing.assert_array_equal(host(vf), v)
        np.testing.assert_array_equal(host(df), d)


@pytest.mark.parametrize("model", [0, 1])
def test_rk_forward_zero_matches_oracle_bitwise(q, model):
    n, ny, Lx, Ly = 64, 64, 3.2, 1.0
    field, coef = O.filtered_noise(n, ny, Lx, Ly, 0.6, 0.2, 0.2, 29)
    rng = np.random.default_rng(3 + model)
    h = rng.uniform(0.2, 0.8, n)
    idx = np.arange(n)
    hl, hr = np.roll(h, 1), np.roll(h, -1)
    dx = Lx / n
    outcomes = set()
    for dtype, c in ((q.F64, coef), (q.F32, coef.astype(np.float32).astype(np.float64))):
        gn = q.FilteredNoise.from_coefficients(coef, dx, Ly / ny, dtype=dtype)
        on = O.OracleNoise.table(c, Lx, Ly)
        # forces spanning blocked sites, sites that find a zero, and sites that
        # march the whole max_advance without one (found = 0)
        for f, max_adv in ((0.0, 1.0), (0.4, 1.0), (3.0, 0.05)):
            rp = q.RkParams(0.5 / ny, max_adv, 1e-11, 1e-9, 3.0)
            rpo = O.rk_params(dstep=0.5 / ny, max_advance=max_adv, root_tol=1e-11,
                              sweep_tol=1e-9, runaway=3.0)
            p = q.Params(Lx, 1.2, f, q.Model(model))
            out, found, evals = q.rk_forward_zero(dev(idx), n, p, gn, rp, dev(h), dev(hl), dev(hr))
            ref = [O.rk_forward_zero(int(i), n, dx, 1.2, f, model, h[i], hl[i], hr[i], on, rpo)
                   for i in idx]
            np.testing.assert_array_equal(host(out), np.array([r[0] for r in ref]))
            np.testing.assert_array_equal(host(found), np.array([r[1] for r in ref]))
            np.testing.assert_array_equal(host(evals), np.array([r[2] for r in ref]))
            outcomes.update((int(r[1]), r[0] == h[i]) for r, i in zip(ref, idx))
    # blocked sites (stay), zeros found after a march, marches without a zero
    assert {(1, True), (1, False), (0, False)} <= outcomes


def test_rk_forward_zero_washboard_and_limits(q):
    # the washboard potential of tests/test_rosso_krauth.cpp:24-31 with no
    # elasticity (lt = 0) below f_c = A k: the site stops where its velocity
    # vanishes; a march of more than 2^26 steps is rejected
    wash = q.SinusoidNoise(0.1, 2 * np.pi)
    A, k = 0.1, 2 * np.pi
    f = 0.5 * A * k
    p = q.Params(16.0, 0.0, f, q.Model.Linear)
    rp = q.RkParams(1 / 200, 1.0, 1e-12, 1e-11, 3.0)
    h0 = np.full(4, 0.3)
    out, found, evals = q.rk_forward_zero(dev(np.arange(4)), 16, p, wash, rp, dev(h0), dev(h0),
                                          dev(h0))
    hz = host(out)
    assert np.all(host(found) == 1) and np.all(hz > 0.3)
    # both edge midpoints sit at hc = (0.3 + h) / 2, where the velocity
    # dx (A k sin(k hc) + f) of the washboard V = A cos(k y) vanishes
    assert np.max(np.abs(A * k * np.sin(k * (0.3 + hz) / 2) + f)) < 1e-9
    with pytest.raises(q.QewInvalidArgument):
        q.rk_forward_zero(dev(np.arange(4)), 16, p, wash, q.RkParams(1e-9, 1.0, 1e-12, 1e-11, 3.0),
                          dev(h0), dev(h0), dev(h0))


@pytest.mark.parametrize("model", [0, 1])
def test_gradient_workspace_edges_match_oracle_bitwise(q, model):
    """The GradientWorkspace's edge_dvdy / edge_slope (src/qew_model.h:126-147)
    as written by the fused kernel equal the oracle's two passes bit for bit,
    and the gradient is the one qew_gradient returns."""
    n, ny = 700, 64  # several tiles, a partial last one
    field, coef = O.filtered_noise(n, ny, 3.5, 1.0, 0.5, 0.2, 0.2, 31)
    on = O.OracleNoise.table(coef, 3.5, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 3.5 / n, 1.0 / ny)
    h = 0.5 + 0.1 * np.cumsum(np.random.default_rng(8).standard_normal(n)) / np.sqrt(n)
    p = q.Params(3.5, 1.3, 0.2, q.Model(model))
    g, e, sl = q.gradient_edges(dev(h), p, gn)
    g_ref, e_ref, s_ref = O.gradient_edges(h, O.params(3.5, 1.3, 0.2, model), on)
    np.testing.assert_array_equal(host(g), g_ref)
    np.testing.assert_array_equal(host(e), e_ref)
    if model == 1:
        np.testing.assert_array_equal(host(sl), s_ref)
    else:
        assert sl is None
    np.testing.assert_array_equal(host(q.gradient(dev(h), p, gn)), g_ref)
