"""GPU parity of the sampler, objective and gradient (HIP path through the C
ABI) against the CPU oracle on the same seeded inputs.

Bar: bit-identical gradients and sampler values for B-spline tables in FP64
(both sides restate the reference's operation order without fused
multiply-adds); energies to 1e-12 relative (device tree sum vs sequential
sum); analytic potentials to 1e-14 (device vs host libm sin/cos).
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, rel, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


def table_pair(q, nx, ny, Lx, Ly, seed, amp=1.0, xi=0.2, dtype=0):
    field, coef = O.filtered_noise(nx, ny, Lx, Ly, amp, xi, xi, seed)
    on = O.OracleNoise.table(coef, Lx, Ly)
    gn = q.FilteredNoise.from_coefficients(coef, Lx / nx, Ly / ny, dtype=dtype)
    return on, gn, coef


def test_sampler_bit_exact(q):
    on, gn, _ = table_pair(q, 48, 40, 2.5, 1.3, 17)
    rng = np.random.default_rng(1)
    x = rng.uniform(-10, 10, 2000)
    y = rng.uniform(-10, 10, 2000)
    ref = O.sample_points(on, x, y)
    got = gn.sample(dev(x), dev(y))
    for r, g in zip(ref, got):
        np.testing.assert_array_equal(host(g), r)


@pytest.mark.parametrize("spacing", ["unit", "odd"])
def test_sampler_bit_exact_at_cell_edges_and_tiny_fractions(q, spacing):
    """Heights and positions on and next to cell boundaries, and tiny
    positive ones whose fractional cell coordinate t makes t^3 / 6 or
    3 t^2 / 6 subnormal: the division-free weight and spacing quotients must
    give the oracle's divisions bit for bit there too (unit spacing, and a
    table whose spacings are not 1)."""
    if spacing == "unit":
        on, gn, _ = table_pair(q, 32, 24, 32.0, 24.0, 5)
        dx, dy = 1.0, 1.0
    else:
        on, gn, _ = table_pair(q, 48, 40, 2.5, 1.3, 17)
        dx, dy = 2.5 / 48, 1.3 / 40
    rng = np.random.default_rng(3)
    k = rng.integers(-60, 60, 300).astype(np.float64)
    ulp = np.spacing(np.abs(k) * dy + 1.0)
    y = np.concatenate([k * dy, k * dy + ulp, k * dy - ulp,
                        [1e-300, 5e-324, 2.0 ** -340 * dy, 2.0 ** -345 * dy, 2.0 ** -520 * dy,
                         2.0 ** -1000, 3.0 * 2.0 ** -1074, 9.0 * 2.0 ** -1074, 0.0, -0.0]])
    extra = rng.uniform(2.0 ** -360, 2.0 ** -330, 200) * dy
    y = np.concatenate([y, extra, -extra])
    x = np.concatenate([rng.integers(-40, 40, len(y) // 2) * dx,
                        rng.uniform(-5, 5, len(y) - len(y) // 2)])
    ref = O.sample_points(on, x, y)
    got = gn.sample(dev(x), dev(y))
    for r, g in zip(ref, got):
        np.testing.assert_array_equal(host(g), r)


def test_sampler_analytic(q):
    on = O.OracleNoise(O.SINCOS, p=(0.3, 0.7, 1.3, 0))
    gn = q.AnalyticNoise(0.3, 0.7, 1.3)
    rng = np.random.default_rng(2)
    x, y = rng.uniform(-5, 5, 500), rng.uniform(-5, 5, 500)
    for r, g in zip(O.sample_points(on, x, y), gn.sample(dev(x), dev(y))):
        np.testing.assert_allclose(host(g), r, atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 254, 255, 256, 257, 510, 511, 1000, 4096])
@pytest.mark.parametrize("model", [0, 1])
def test_gradient_and_energy_match_oracle(q, n, model):
    nx, ny = max(n, 8), 64
    on, gn, _ = table_pair(q, nx, ny, float(nx) * 0.05, 1.0, 3 + n)
    p_o = O.params(float(n) * 0.05 * (nx / max(n, 8)), 0.8, 0.13, model)
    p = q.Params(p_o.physical_size, 0.8, 0.13, q.Model(model))
    h = np.random.default_rng(n).uniform(-2, 3, n)
    g_ref = O.gradient(h, p_o, on)
    e_ref = O.objective(h, p_o, on)
    hd = dev(h)
    np.testing.assert_array_equal(host(q.gradient(hd, p, gn)), g_ref)
    e, g = q.evaluate(hd, p, gn)
    np.testing.assert_array_equal(host(g), g_ref)
    assert rel(e, e_ref) < 1e-12 or abs(e - e_ref) < 1e-14 * n
    assert rel(q.objective(hd, p, gn), e_ref) < 1e-12 or abs(q.objective(hd, p, gn) - e_ref) < 1e-14 * n


def test_objective_trial_matches_oracle(q):
    on, gn, _ = table_pair(q, 256, 64, 2.56, 1.0, 9)
    p_o = O.params(2.56, 2.0, 0.1, 1)
    p = q.Params(2.56, 2.0, 0.1, q.Model.Arclength)
    rng = np.random.default_rng(4)
    x0, d = rng.uniform(0, 1, 256), rng.standard_normal(256)
    e_ref = O.objective_trial(x0, d, 0.37, p_o, on)
    assert rel(q.objective_trial(dev(x0), dev(d), 0.37, p, gn), e_ref) < 1e-12


@pytest.mark.parametrize("kind", ["sincos", "const", "zero", "harmonic", "washboard"])
def test_analytic_potentials(q, kind):
    spec = {"sincos": (O.SINCOS, q.AnalyticNoise, (0.3, 0.7, 1.3)),
            "const": (O.CONST, q.ConstNoise, (1.25,)),
            "zero": (O.ZERO, q.ZeroNoise, ()),
            "harmonic": (O.HARMONIC, q.HarmonicWell, (2.0, 0.5)),
            "washboard": (O.WASHBOARD, q.SinusoidNoise, (0.1, 2 * np.pi))}[kind]
    on = O.OracleNoise(spec[0], p=tuple(spec[2]) + (0,) * (4 - len(spec[2])))
    gn = spec[1](*spec[2])
    h = np.random.default_rng(5).uniform(-1, 1, 300)
    for model in (0, 1):
        p_o = O.params(300.0, 1.1, 0.37, model)
        p = q.Params(300.0, 1.1, 0.37, q.Model(model))
        np.testing.assert_allclose(host(q.gradient(dev(h), p, gn)), O.gradient(h, p_o, on),
                                   rtol=1e-13, atol=1e-14)
        e = q.objective(dev(h), p, gn)
        assert abs(e - O.objective(h, p_o, on)) < 1e-12 * max(1, abs(e))


def test_flat_line_constant_potential_closed_form(q):
    # reference tests/test_qew_model.cpp:110-132
    for model in (q.Model.Linear, q.Model.Arclength):
        e = q.objective(dev(np.full(50, 0.7)), q.Params(3.0, 2.3, 0.4, model), q.ConstNoise(1.25))
        assert abs(e - 3.0 * (1.25 - 0.4 * 0.7)) < 1e-12


def test_table_fp32_storage_close(q):
    on, gn32, _ = table_pair(q, 512, 128, 5.12, 1.0, 21, dtype=1)
    p_o = O.params(5.12, 1.0, 0.0, 0)
    p = q.Params(5.12, 1.0, 0.0, q.Model.Linear)
    h = np.random.default_rng(6).uniform(0, 1, 512)
    g_ref = O.gradient(h, p_o, on)
    g = host(q.gradient(dev(h), p, gn32))
    # FP32 coefficient storage: ~6e-8 relative per coefficient
    assert np.max(np.abs(g - g_ref)) < 1e-5 * np.max(np.abs(g_ref))


def test_noise_generation_matches_oracle(q):
    for args in [(64, 48, 3.2, 1.0, 1.5, 0.15, 0.15, 3), (48, 40, 2.5, 1.3, 1.0, 0.25, 0.25, 17),
                 (100, 30, 4.0, 1.0, 2.0, 0.2, 0.1, 42)]:
        field, coef = O.filtered_noise(*args)
        gn = q.FilteredNoise(*args)
        np.testing.assert_allclose(gn.field(), field, atol=1e-12 * args[4])
        np.testing.assert_allclose(gn.coefficients(), coef, atol=1e-11 * args[4])


def test_noise_from_field_golden(q):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "noise_golden.json")))
    nx, ny = int(g["nx"]), int(g["ny"])
    field = np.array(g["field"]["data"]).reshape(nx, ny)
    gn = q.FilteredNoise.from_field(field, g["Lx"], g["Ly"])
    v, dvx, dvy = gn.sample(dev(np.array(g["x"]["data"])), dev(np.array(g["y"]["data"])))
    assert np.max(np.abs(host(v) - g["v"]["data"])) < 1e-7
    assert np.max(np.abs(host(dvx) - g["dvdx"]["data"])) < 1e-4
    assert np.max(np.abs(host(dvy) - g["dvdy"]["data"])) < 1e-4


def test_batch_eval_ragged_equals_single_lines(q):
    torch = torch_cuda()
    # tile boundaries at multiples of 255 sites: 255, 256, 510, 511
    lengths = [1024, 77, 2048, 256, 4096, 3, 255, 511, 510]
    lines, noises, params, oracle = [], [], [], []
    for k, n in enumerate(lengths):
        on, gn, _ = table_pair(q, max(n, 8), 32, max(n, 8) * 1.0, 32.0, 100 + k, xi=4.0)
        p = q.Params(float(max(n, 8)), 1.0 + 0.1 * k, 0.05 * k, q.Model(k % 2))
        noises.append(gn)
        params.append(p)
        oracle.append((on, O.params(p.physical_size, p.line_tension, p.driving_force, k % 2)))
        lines.append(np.random.default_rng(k).uniform(0, 32, n))
    b = q.Batch(lengths, noises, params)
    h = dev(np.concatenate(lines))
    e, g = b.eval(h)
    g = host(g)
    e = host(e)
    for k, n in enumerate(lengths):
        on, po = oracle[k]
        sl = slice(b.offsets[k], b.offsets[k + 1])
        np.testing.assert_array_equal(g[sl], O.gradient(lines[k], po, on))
        er = O.objective(lines[k], po, on)
        assert abs(e[k] - er) < 1e-12 * max(1.0, abs(er))


def test_batch_eval_fp32_compensated_error(q):
    """FP32 line state + FP32 table + compensated energy vs the FP64 oracle.
    Stated tolerance: gradient max error < 2e-4 of max|grad|, energy < 1e-5
    relative."""
    torch = torch_cuda()
    n, ny = 8192, 512
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 8.0, 8.0, 1)
    on = O.OracleNoise.table(coef, float(n), float(ny))
    gn32 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0, dtype=q.F32)
    rng = np.random.default_rng(12345)
    h = ny / 2 + np.cumsum(rng.standard_normal(n))
    p_o = O.params(float(n), 1.0, 0.0, 1)
    b = q.Batch([n], [gn32], [q.Params(float(n), 1.0, 0.0, q.Model.Arclength)])
    anchor = dev(np.array([ny / 2]))
    off = dev((h - ny / 2).astype(np.float32))
    e, g = b.eval(off, dtype=q.F32, anchor=anchor)
    g_ref = O.gradient(h, p_o, on)
    e_ref = O.objective(h, p_o, on)
    assert np.max(np.abs(host(g) - g_ref)) < 2e-4 * np.max(np.abs(g_ref))
    assert rel(host(e)[0], e_ref) < 1e-5


def test_noise_generate_batch_equals_single_and_convert(q):
    specs = [(64, 48, 3.2, 1.0, 1.5, 0.15, 0.15, 3), (48, 40, 2.5, 1.3, 1.0, 0.25, 0.25, 17),
             (100, 30, 4.0, 1.0, 2.0, 0.2, 0.1, 42), (32, 32, 32.0, 32.0, 1.0, 4.0, 4.0, 9)]
    batch = q.FilteredNoise.generate_batch(specs)
    for spec, nb in zip(specs, batch):
        single = q.FilteredNoise(*spec)
        np.testing.assert_array_equal(nb.coefficients(), single.coefficients())
        np.testing.assert_array_equal(nb.field(), single.field())
        c32 = nb.convert(q.F32).coefficients()
        np.testing.assert_array_equal(c32, nb.coefficients().astype(np.float32).astype(np.float64))
