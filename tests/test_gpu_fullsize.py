"""Parity at the full size of the gradient/objective throughput configuration:
L = 65536 sites on a 65536 x 8192 FilteredNoise table (xi = 8 grid units,
seed 1), arclength elasticity, a seeded unit-variance random walk started at
Nh/2 -- the workload bench.py times. The table is generated on the GPU and
its coefficients are handed to the oracle, so both sides see the same table.

Bar: FP64 gradient bit-identical to the oracle, energy within 1e-12
relative; FP32-compensated batch evaluation within the stated error
(gradient < 2e-4 of max|grad|, energy < 1e-5 relative); evaluation r of a
batch of shifted walks equals the single evaluation of that walk bit for
bit. Host memory: one 4.3 GB copy of the table.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, rel, torch_cuda

pytestmark = pytest.mark.gpu

L, NY, XI = 65536, 8192, 8.0


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


def test_full_size_eval_matches_oracle(q):
    noise = q.FilteredNoise(L, NY, float(L), float(NY), 1.0, XI, XI, 1, dtype=q.F64)
    coef = noise.coefficients()
    on = O.OracleNoise.table(coef, float(L), float(NY))
    h = NY / 2 + np.cumsum(np.random.default_rng(2).standard_normal(L))
    p_o = O.params(float(L), 1.0, 0.0, O.ARCLENGTH)
    p = q.Params(float(L), 1.0, 0.0, q.Model.Arclength)
    g_ref = O.gradient(h, p_o, on, omp=True)
    e_ref = O.objective(h, p_o, on, omp=True)

    e, g = q.evaluate(dev(h), p, noise)
    np.testing.assert_array_equal(host(g), g_ref)
    assert rel(e, e_ref) < 1e-12

    # a batch of three walks shifted by 8 cells (the benchmark's evaluations)
    shifts = np.array([0.0, 8.0, 16.0])
    b = q.Batch([L] * 3, [noise] * 3, [p] * 3)
    eb, gb = b.eval(dev(np.concatenate([h + s for s in shifts])))
    gb, eb = host(gb), host(eb)
    np.testing.assert_array_equal(gb[:L], g_ref)
    for k, s in enumerate(shifts):
        e1, g1 = q.evaluate(dev(h + s), p, noise)
        np.testing.assert_array_equal(gb[k * L:(k + 1) * L], host(g1))
        assert eb[k] == e1

    # FP32 line offsets from an anchor, FP32 table, compensated energy
    n32 = noise.convert(q.F32)
    b32 = q.Batch([L], [n32], [p])
    off = dev((h - NY / 2).astype(np.float32))
    e32, g32 = b32.eval(off, dtype=q.F32, anchor=dev(np.array([NY / 2])))
    assert np.max(np.abs(host(g32) - g_ref)) < 2e-4 * np.max(np.abs(g_ref))
    assert rel(host(e32)[0], e_ref) < 1e-5


def test_full_size_langevin_matches_oracle(q):
    """The Langevin anneal configuration (L = 65536 on a 65536 x 4096 table,
    xi = 4, seed 1, linear, line tension 1, f = 0.256, dt = 0.01 at T = 0.5,
    Philox key 2), 20 steps through the batch entry point (several steps per
    launch) against 20 oracle steps: agreement to a few ulp per step (device
    vs host libm in Box-Muller), as for the small lines; FP32 state from the
    same normals within 1e-5 of the FP64 heights' mean and spread."""
    ny = 4096
    noise = q.FilteredNoise(L, ny, float(L), float(ny), 1.0, 4.0, 4.0, 1, dtype=q.F64)
    on = O.OracleNoise.table(noise.coefficients(), float(L), float(ny))
    p_o = O.params(float(L), 1.0, 0.256, O.LINEAR)
    p = q.Params(float(L), 1.0, 0.256, q.Model.Linear)
    steps, dt, temp = 20, 0.01, 0.5
    drift = np.full(steps, dt)
    diff = np.full(steps, np.sqrt(2.0 * temp * dt))
    h0 = ny / 2 + np.random.default_rng(7).uniform(-0.5, 0.5, L)
    h_ref = h0.copy()
    for s in range(steps):
        O.langevin_step(h_ref, p_o, on, dt, diff[s], 2, s)
    b = q.Batch([L], [noise], [p])
    hb = dev(h0)
    b.langevin(hb, drift, diff, seed=2)
    hg = host(hb)
    assert np.max(np.abs(hg - h_ref)) < 1e-11

    b32 = q.Batch([L], [noise.convert(q.F32)], [p])
    off = dev((h0 - ny / 2).astype(np.float32))
    b32.langevin(off, drift, diff, seed=2, dtype=q.F32, anchor=dev(np.array([ny / 2])))
    h32 = host(off).astype(np.float64) + ny / 2
    assert rel(np.mean(h32 - ny / 2), np.mean(h_ref - ny / 2)) < 1e-5
    assert rel(np.var(h32), np.var(h_ref)) < 1e-5
