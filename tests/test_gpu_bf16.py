"""BF16 coefficient tables (storage only; FP32 line state and arithmetic).

The conversion rounds FP64 -> FP32 -> BF16 to nearest even; a BF16 table
read by the FP32 batch evaluation and dynamics gives bit for bit what an FP32
table holding the same (BF16-representable) values gives, so the only error
is the coefficients' rounding, stated against the FP64 oracle below. Every
other entry point rejects BF16 tables.
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


def bf16_round(x):
    """FP64 -> FP32 -> BF16 (round to nearest even) -> FP64, in numpy."""
    u = np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_conversion_rounds_to_nearest_even(q):
    # odd and even nx, a table with values that exercise ties and carries
    for nx, ny, seed in [(64, 48, 3), (67, 33, 5)]:
        field, coef = O.filtered_noise(nx, ny, float(nx), float(ny), 1.0, 4.0, 4.0, seed)
        g = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
        b = g.convert(q.BF16)
        assert b.info()["dtype"] == q.BF16
        np.testing.assert_array_equal(b.coefficients(), bf16_round(coef))
    # the rounding helper itself on known values (bfloat16 has 8 significand bits)
    assert bf16_round(1.0 + 2.0 ** -8) == 1.0                   # tie -> even
    assert bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6    # tie -> even (up)
    assert bf16_round(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7        # representable


def _tables(q, n, ny, seed):
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, seed)
    g64 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
    g16 = g64.convert(q.BF16)
    g32r = q.FilteredNoise.from_coefficients(bf16_round(coef), 1.0, 1.0, dtype=q.F32)
    return coef, g64, g16, g32r


def test_bf16_batch_eval_equals_fp32_of_rounded_values(q):
    # lines with odd and even window starts, the wrap edge, ragged lengths
    n, ny = 1026, 128
    coef, g64, g16, g32r = _tables(q, n, ny, 7)
    lengths = [n, n // 2, 2 * n]
    models = [q.Model.Arclength, q.Model.Linear, q.Model.Arclength]
    params = [q.Params(float(n), 1.0, 0.02, m) for m in models]
    rng = np.random.default_rng(4)
    offs = [np.cumsum(rng.standard_normal(L)).astype(np.float32) for L in lengths]
    h = dev(np.concatenate(offs))
    anchor = dev(np.array([ny / 2.0, ny / 2.0 + 0.3, 3.7]))
    e16, g16v = q.Batch(lengths, [g16] * 3, params).eval(h, dtype=q.F32, anchor=anchor)
    e32, g32v = q.Batch(lengths, [g32r] * 3, params).eval(h, dtype=q.F32, anchor=anchor)
    np.testing.assert_array_equal(host(g16v), host(g32v))
    np.testing.assert_array_equal(host(e16), host(e32))


def test_bf16_batch_langevin_equals_fp32_of_rounded_values(q):
    n, ny = 700, 128
    coef, g64, g16, g32r = _tables(q, n, ny, 9)
    p = q.Params(float(n), 1.0, 0.05, q.Model.Linear)
    o0 = np.random.default_rng(2).normal(0, 1, n).astype(np.float32)
    anchor = dev(np.array([ny / 2.0]))
    steps = 37
    drift = np.full(steps, 0.01)
    diff = np.full(steps, 0.03)
    a, b = dev(o0), dev(o0)
    q.Batch([n], [g16], [p]).langevin(a, drift, diff, seed=5, dtype=q.F32, anchor=anchor)
    q.Batch([n], [g32r], [p]).langevin(b, drift, diff, seed=5, dtype=q.F32, anchor=anchor)
    np.testing.assert_array_equal(host(a), host(b))


def test_bf16_error_against_fp64(q):
    """Stated BF16 error on a unit-variance table (xi = 8, arclength random
    walk): gradient max error < 2e-2 of max|grad|, energy < 1e-3 relative."""
    n, ny = 8192, 512
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 8.0, 8.0, 1)
    on = O.OracleNoise.table(coef, float(n), float(ny))
    g16 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0).convert(q.BF16)
    h = ny / 2 + np.cumsum(np.random.default_rng(12345).standard_normal(n))
    b = q.Batch([n], [g16], [q.Params(float(n), 1.0, 0.0, q.Model.Arclength)])
    e, g = b.eval(dev((h - ny / 2).astype(np.float32)), dtype=q.F32, anchor=dev(np.array([ny / 2])))
    p_o = O.params(float(n), 1.0, 0.0, 1)
    g_ref = O.gradient(h, p_o, on)
    e_ref = O.objective(h, p_o, on)
    assert np.max(np.abs(host(g) - g_ref)) < 2e-2 * np.max(np.abs(g_ref))
    assert rel(host(e)[0], e_ref) < 1e-3


def test_bf16_tables_rejected_by_other_entry_points(q):
    n, ny = 64, 32
    coef, g64, g16, g32r = _tables(q, n, ny, 11)
    p = q.Params(float(n), 1.0, 0.0, q.Model.Linear)
    h = dev(np.full(n, ny / 2.0))
    with pytest.raises(q.QewInvalidArgument):
        q.gradient(h, p, g16)
    with pytest.raises(q.QewInvalidArgument):
        q.objective(h, p, g16)
    with pytest.raises(q.QewInvalidArgument):
        q.lbfgs_minimize(h, p, g16)
    with pytest.raises(q.QewInvalidArgument):
        q.rk_block(h, p, g16, q.RkParams(0.5, 1.0, 1e-10, 1e-9, 2.0))
    with pytest.raises(q.QewInvalidArgument):
        q.LangevinDynamics(n, 1).step(h, p, g16, 0.01, 0.0)
    with pytest.raises(q.QewInvalidArgument):
        g16.sample(dev(np.array([1.0])), dev(np.array([1.0])))
    with pytest.raises(q.QewInvalidArgument):
        g16.convert(q.F32)
    b = q.Batch([n], [g16], [p])
    with pytest.raises(q.QewInvalidArgument):
        b.eval(h)                       # FP64 line state
    with pytest.raises(q.QewInvalidArgument):
        b.lbfgs(h, q.LbfgsParams(m=4, precondition=False))
    with pytest.raises(q.QewInvalidArgument):
        b.rk_block(h, q.RkParams(0.5, 1.0, 1e-10, 1e-9, 2.0))
    with pytest.raises(q.QewInvalidArgument):
        b.langevin(h, [0.01], [0.0], seed=1)   # FP64 line state
