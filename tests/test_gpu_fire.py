"""GPU checks of FIRE (src/fire.h:41-131), single line and batched, through
the C ABI.

The reference's own FIRE tests (tests/test_fire.cpp: flat line, harmonic
minimum, energy decrease, determinism, non-convergence) and its L-BFGS/FIRE
agreement test (tests/test_lbfgs.cpp:246-271) are restated on the GPU path.
Against the CPU oracle's sequential FIRE the device runs the same element
updates with workgroup tree sums for P, |v|^2 and |F|^2, so trajectories
agree to rounding: after a few steps heights within 1e-12, at convergence
(ftol 1e-9) heights within 1e-9, iteration counts within 2, energies within
1e-12 relative.
"""
import math

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


def lin(n, model=0):
    import qew_amd
    return qew_amd.Params(float(n), 1.0, 0.0, qew_amd.Model(model))


def fire_params(q, **kw):
    return q.FireParams(dt_max=0.1 * math.sqrt(1.0), **kw)


def test_relaxes_to_flat_line(q):
    n = 128
    h = dev(0.5 + 0.3 * np.cos(4 * np.pi * np.arange(n) / n))
    r = q.fire_minimize(h, lin(n), q.ZeroNoise(), fire_params(q, ftol=1e-9))
    assert r.converged and r.max_force < 1e-9
    assert np.ptp(host(h)) < 1e-6


@pytest.mark.parametrize("model", [0, 1])
def test_converges_to_harmonic_minimum(q, model):
    h = dev(np.zeros(96))
    r = q.fire_minimize(h, lin(96, model), q.HarmonicWell(2.0, 0.5), fire_params(q, ftol=1e-9))
    assert r.converged and np.max(np.abs(host(h) - 0.5)) < 1e-6


def test_lowers_energy(q):
    rng = np.random.default_rng(1)
    h0 = rng.random(96)
    p, well = lin(96, 1), q.HarmonicWell(1.5, 0.6)
    h = dev(h0)
    e0 = q.objective(h, p, well)
    q.fire_minimize(h, p, well, fire_params(q, ftol=1e-9))
    assert q.objective(h, p, well) <= e0 + 1e-12


def test_deterministic(q):
    n = 64
    h0 = 0.3 * np.sin(2 * np.pi * np.arange(n) / n)
    p, well = lin(n), q.HarmonicWell(1.0, 0.5)
    h1, h2 = dev(h0), dev(h0)
    r1 = q.fire_minimize(h1, p, well, fire_params(q, ftol=1e-10))
    r2 = q.fire_minimize(h2, p, well, fire_params(q, ftol=1e-10))
    assert r1.iterations == r2.iterations
    assert np.array_equal(host(h1), host(h2))


def test_reports_nonconvergence(q):
    h = dev(np.zeros(64))
    r = q.fire_minimize(h, lin(64), q.HarmonicWell(2.0, 0.5),
                        fire_params(q, ftol=1e-12, max_iter=3))
    assert not r.converged and r.iterations == 3 and r.max_force >= 1e-12
    h0 = dev(np.zeros(8))
    r0 = q.fire_minimize(h0, lin(8), q.HarmonicWell(2.0, 0.5), fire_params(q, max_iter=0))
    assert r0.iterations == 0 and not r0.converged and r0.max_force > 0
    assert np.array_equal(host(h0), np.zeros(8))


def test_agrees_with_lbfgs(q):
    rng = np.random.default_rng(3)
    h0 = rng.random(96)
    p, well = lin(96), q.HarmonicWell(1.5, 0.4)
    ha, hb = dev(h0), dev(h0)
    q.lbfgs_minimize(ha, p, well, q.LbfgsParams(ftol=1e-9))
    r = q.fire_minimize(hb, p, well, fire_params(q, ftol=1e-9))
    assert r.converged and np.max(np.abs(host(ha) - host(hb))) < 1e-5


def test_rejects_missing_dt_max(q):
    h = dev(np.zeros(8))
    with pytest.raises(q.QewInvalidArgument):
        q.fire_minimize(h, lin(8), q.ZeroNoise(), q.FireParams())


def _table_problem(n, ny, xi, seed):
    _, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, xi, xi, seed)
    return coef


@pytest.mark.parametrize("model", [0, 1])
def test_matches_oracle_on_table(q, model):
    n, ny = 128, 64
    coef = _table_problem(n, ny, 4.0, 11 + model)
    nz = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
    onz = O.OracleNoise.table(coef, float(n), float(ny))
    p = q.Params(float(n), 2.0, 0.0, q.Model(model))
    op = O.params(float(n), 2.0, 0.0, model)
    rng = np.random.default_rng(model)
    h0 = ny / 2.0 + rng.standard_normal(n)
    # a few steps: the same trajectory to rounding
    h = dev(h0)
    r = q.fire_minimize(h, p, nz, q.FireParams(dt_max=0.2, ftol=0.0, max_iter=5))
    ho, ro = O.fire(h0, op, onz, O.FireParams(0.2, ftol=0.0, max_iter=5))
    assert r.iterations == ro.iterations == 5
    assert np.max(np.abs(host(h) - ho)) < 1e-12
    # to convergence: the same minimum
    h = dev(h0)
    r = q.fire_minimize(h, p, nz, q.FireParams(dt_max=0.2, ftol=1e-9))
    ho, ro = O.fire(h0, op, onz, O.FireParams(0.2, ftol=1e-9))
    assert r.converged and ro.converged
    assert abs(r.iterations - ro.iterations) <= 2
    assert np.max(np.abs(host(h) - ho)) < 1e-9
    assert rel(O.objective(host(h), op, onz), O.objective(ho, op, onz)) < 1e-12


def test_batch_matches_single_lines_and_oracle(q):
    lengths = [96, 33, 256, 1, 200]
    models = [0, 1, 1, 0, 0]
    ny = 64
    noises, params, oracle, h0 = [], [], [], []
    for k, (n, model) in enumerate(zip(lengths, models)):
        coef = _table_problem(max(n, 4), ny, 3.0, 40 + k)
        noises.append(q.FilteredNoise.from_coefficients(coef, 1.0, 1.0))
        params.append(q.Params(float(max(n, 4)), 1.5, 0.01, q.Model(model)))
        oracle.append((O.OracleNoise.table(coef, float(max(n, 4)), float(ny)),
                       O.params(float(max(n, 4)), 1.5, 0.01, model)))
        h0.append(ny / 2.0 + 0.5 * np.random.default_rng(k).standard_normal(n))
    fp = q.FireParams(dt_max=0.2, ftol=1e-9, max_iter=20000)
    b = q.Batch(lengths, noises, params)
    hb = dev(np.concatenate(h0))
    res = b.fire(hb, fp)
    hb = host(hb)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    for k, n in enumerate(lengths):
        hs = dev(h0[k])
        rs = q.fire_minimize(hs, params[k], noises[k], fp)
        assert (res[k].iterations, res[k].converged) == (rs.iterations, rs.converged)
        assert np.array_equal(hb[offs[k]:offs[k + 1]], host(hs))
        ho, ro = O.fire(h0[k], oracle[k][1], oracle[k][0],
                        O.FireParams(0.2, ftol=1e-9, max_iter=20000))
        assert res[k].converged == bool(ro.converged)
        assert abs(res[k].iterations - ro.iterations) <= 2
        assert np.max(np.abs(hb[offs[k]:offs[k + 1]] - ho)) < 1e-9


def test_fp32_table_matches_oracle_on_rounded_coefficients(q):
    n, ny = 128, 64
    coef = _table_problem(n, ny, 4.0, 77)
    nz32 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0, dtype=q.F32)
    onz = O.OracleNoise.table(coef.astype(np.float32).astype(np.float64), float(n), float(ny))
    p, op = q.Params(float(n), 2.0, 0.0, q.Model(0)), O.params(float(n), 2.0, 0.0, 0)
    h0 = np.full(n, ny / 2.0)
    b = q.Batch([n], [nz32], [p])
    h = dev(h0)
    r = b.fire(h, q.FireParams(dt_max=0.2, ftol=1e-9))[0]
    ho, ro = O.fire(h0, op, onz, O.FireParams(0.2, ftol=1e-9))
    assert r.converged and ro.converged and abs(r.iterations - ro.iterations) <= 2
    assert np.max(np.abs(host(h) - ho)) < 1e-9
