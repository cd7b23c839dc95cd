"""GPU checks of the persistent batched L-BFGS (one workgroup per line,
device-side control flow) against the CPU oracle's sequential lbfgs_minimize.

FP64 mode: same algorithm and arithmetic as the oracle, with workgroup tree
sums instead of sequential sums, so every converged line must reach the
oracle's minimum: energy within 1e-11 relative, heights within 1e-5 (ftol
1e-6: above the Armijo roundoff floor of these grid-unit lines, where both
sides converge).
FP32 mode (FP32 line offsets and tables, Kahan dots, FP64 Armijo and curvature
pair): stated tolerance 1e-6 relative energy against the FP64 oracle minimum.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, rel, torch_cuda

pytestmark = pytest.mark.gpu
LT = 8.0  # line tension: lambda_p / xi = 4 at unit amplitude


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


def make_problem(q, lengths, ny, xis, models, seeds, dtype):
    noises, oracle, params, h0 = [], [], [], []
    for n, xi, model, seed in zip(lengths, xis, models, seeds):
        field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, xi, xi, seed)
        oracle.append((O.OracleNoise.table(coef, float(n), float(ny)),
                       O.params(float(n), LT, 0.0, model)))
        noises.append(q.FilteredNoise.from_coefficients(coef, 1.0, 1.0, dtype=dtype))
        params.append(q.Params(float(n), LT, 0.0, q.Model(model)))
        h0.append(np.full(n, ny / 2.0))
    return noises, oracle, params, h0


@pytest.mark.parametrize("precondition", [True, False])
def test_batch_lbfgs_fp64_matches_oracle(q, precondition):
    lengths = [256, 64, 512, 128]
    xis = [2.0, 4.0, 8.0, 4.0]
    models = [0, 1, 0, 1]
    noises, oracle, params, h0 = make_problem(q, lengths, 128, xis, models, [1, 2, 3, 4], q.F64)
    lp_kw = dict(m=10, ftol=1e-6, max_iter=4000, precondition=precondition)
    b = q.Batch(lengths, noises, params)
    h = dev(np.concatenate(h0))
    res = b.lbfgs(h, q.LbfgsParams(**lp_kw, precond_shift=0.5 / 16.0), dtype=q.F64)
    hh = host(h)
    for k, n in enumerate(lengths):
        on, po = oracle[k]
        h_ref, r_ref = O.lbfgs(h0[k], po, on, O.lbfgs_params(**lp_kw, precond_shift=0.5 / 16.0))
        assert res[k].converged and r_ref.converged, (k, res[k], r_ref.iterations)
        line = hh[b.offsets[k]:b.offsets[k + 1]]
        e = O.objective(line, po, on)
        e_ref = O.objective(h_ref, po, on)
        assert rel(e, e_ref) < 1e-11, (k, e, e_ref)
        assert np.max(np.abs(line - h_ref)) < 1e-5


@pytest.mark.parametrize("iters", [1, 2])
@pytest.mark.parametrize("lengths", [[16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384],
                                     [14, 96, 640, 1000, 256]], ids=["pow2", "mixed"])
def test_batch_lbfgs_first_steps_match_oracle(q, iters, lengths):
    """The first one or two preconditioned steps against the oracle's: the
    step is alpha P^-1 g with P applied through the LDS FFT, so every radix
    sequence of the transform (n/2 = 8 ... 8192 complex points on the
    power-of-two kernel: radix 8 alone, 8 then 2, 8 then 4, ...; and the
    mixed-radix kernel, which a batch with any other even length uses for all
    its lines) is checked to rounding, far below what a wrong butterfly or
    twiddle would give."""
    ny = 128
    xis = [2.0 + (k % 3) for k in range(len(lengths))]
    models = [k % 2 for k in range(len(lengths))]
    noises, oracle, params, h0 = make_problem(q, lengths, ny, xis, models,
                                              list(range(31, 31 + len(lengths))), q.F64)
    lp_kw = dict(m=10, ftol=1e-12, max_iter=iters, precondition=True)
    b = q.Batch(lengths, noises, params)
    h = dev(np.concatenate(h0))
    res = b.lbfgs(h, q.LbfgsParams(**lp_kw, precond_shift=0.5 / 16.0), dtype=q.F64)
    hh = host(h)
    for k, n in enumerate(lengths):
        on, po = oracle[k]
        h_ref, r_ref = O.lbfgs(h0[k], po, on, O.lbfgs_params(**lp_kw, precond_shift=0.5 / 16.0),
                               omp=False)
        assert res[k].iterations == r_ref.iterations == iters, (k, n, res[k], r_ref.iterations)
        line = hh[b.offsets[k]:b.offsets[k + 1]]
        moved = np.max(np.abs(h_ref - h0[k]))
        assert moved > 1e-3, (k, n, moved)
        assert np.max(np.abs(line - h_ref)) < 1e-11 * max(1.0, moved), (k, n)


def test_batch_lbfgs_fp32_energy_error(q):
    lengths = [1024, 2048, 512, 1024, 4096, 256]
    xis = [2.0, 4.0, 8.0, 4.0, 8.0, 2.0]
    models = [0, 1, 0, 1, 1, 0]
    noises, oracle, params, h0 = make_problem(q, lengths, 256, xis, models, [5, 6, 7, 8, 9, 10], q.F32)
    b = q.Batch(lengths, noises, params)
    h = dev(np.concatenate(h0))
    res = b.lbfgs(h, q.LbfgsParams(m=10, ftol=1e-4, max_iter=4000, precondition=True,
                                   precond_shift=0.5 / 16.0), dtype=q.F32)
    hh = host(h)
    for k, n in enumerate(lengths):
        on, po = oracle[k]
        h_ref, r_ref = O.lbfgs(h0[k], po, on, O.lbfgs_params(m=10, ftol=1e-6, max_iter=4000,
                                                             precondition=True,
                                                             precond_shift=0.5 / 16.0))
        line = hh[b.offsets[k]:b.offsets[k + 1]]
        e = O.objective(line, po, on)
        e_ref = O.objective(h_ref, po, on)
        assert rel(e, e_ref) < 1e-6, (k, res[k], e, e_ref)


def test_batch_lbfgs_harmonic_and_branches(q):
    # tests/test_lbfgs.cpp:87-188 restated on the batched kernel (analytic
    # potentials, FP64)
    well = q.HarmonicWell(2.0, 0.5)
    p = q.Params(96.0, 1.0, 0.0, q.Model.Arclength)
    b = q.Batch([96, 96], [well, well], [p, q.Params(96.0, 1.0, 0.0, q.Model.Linear)])
    h = dev(np.zeros(192))
    res = b.lbfgs(h, q.LbfgsParams(ftol=1e-9))
    assert all(r.converged for r in res)
    assert np.max(np.abs(host(h) - 0.5)) < 1e-6
    # max_ls = 0: gives up at iteration 0 without history
    h = dev(np.zeros(192))
    res = b.lbfgs(h, q.LbfgsParams(ftol=1e-9, max_ls=0))
    assert all((not r.converged) and r.iterations == 0 and r.n_ls_restarts == 0 for r in res)
    # curvature pairs all rejected: still converges
    h = dev(np.zeros(192))
    b16 = q.Batch([16, 16], [well, well], [q.Params(16.0, 1.0, 0.0, q.Model.Linear)] * 2)
    h = dev(np.zeros(32))
    res = b16.lbfgs(h, q.LbfgsParams(ftol=1e-8, max_iter=20000, curvature_eps=1e30))
    assert all(r.converged and r.n_curvature_skips > 0 for r in res)


def test_batch_lbfgs_rejects_bad_lengths_for_fft(q):
    # the LDS preconditioner takes even lengths whose half has no prime factor
    # above 7: 22 = 2 * 11 and the odd 105 are rejected
    well = q.HarmonicWell(2.0, 0.5)
    for n in (22, 105):
        b = q.Batch([n], [well], [q.Params(float(n), 1.0, 0.0, q.Model.Linear)])
        with pytest.raises(ValueError):
            b.lbfgs(dev(np.zeros(n)), q.LbfgsParams(precondition=True))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_lbfgs_mixed_radix_preconditioner_matches_oracle(q, dtype):
    """Preconditioned lines whose half-length has factors 3, 5 and 7 (the
    reference preconditions any n, src/lbfgs.h:116-146; n = 100 is the
    length of its FFT test, tests/test_fft_util.cpp:9) in one batch with
    power-of-two ones: the FP64 kernel reaches the oracle's minima (energy
    1e-11 relative) in about as many iterations, FP32 within its stated 1e-6."""
    lengths = [100, 96, 120, 126, 1000, 6000, 14, 256]
    xis = [2.0, 4.0, 4.0, 2.0, 8.0, 4.0, 2.0, 4.0]
    models = [0, 1, 0, 1, 0, 1, 0, 1]
    noises, oracle, params, h0 = make_problem(q, lengths, 128, xis, models,
                                              list(range(31, 31 + len(lengths))), q.F64)
    shifts = [0.5 / xi ** 2 for xi in xis]
    kw = dict(m=10, ftol=1e-6, max_iter=4000, precondition=True)
    if dtype == "f32":
        noises = [nz.convert(q.F32) for nz in noises]
    b = q.Batch(lengths, noises, params)
    b.set_precond_shifts(shifts)
    h = dev(np.concatenate(h0))
    res = b.lbfgs(h, q.LbfgsParams(**{**kw, "ftol": 1e-6 if dtype == "f64" else 1e-4}),
                  dtype=q.F64 if dtype == "f64" else q.F32)
    hh = host(h)
    for k, n in enumerate(lengths):
        on, po = oracle[k]
        h_ref, r_ref = O.lbfgs(h0[k], po, on, O.lbfgs_params(**kw, precond_shift=shifts[k]))
        assert r_ref.converged and res[k].converged, (k, n, res[k], r_ref.iterations)
        e = O.objective(hh[b.offsets[k]:b.offsets[k + 1]], po, on)
        e_ref = O.objective(h_ref, po, on)
        if dtype == "f64":
            assert rel(e, e_ref) < 1e-11, (k, n, e, e_ref)
            assert abs(res[k].iterations - r_ref.iterations) <= max(5, r_ref.iterations // 5), \
                (k, n, res[k].iterations, r_ref.iterations)
        else:
            assert rel(e, e_ref) < 1e-6, (k, n, e, e_ref)


def test_batch_lbfgs_long_lines_match_oracle(q):
    """Config-4 line lengths (2048 .. 16384, tables L x 2048): the persistent
    kernel's multi-butterfly LDS FFT stages and long reductions. FP64 must
    reach the oracle's minimum (energy 1e-10 relative) in a similar number of
    iterations; FP32 within the stated 1e-6 relative energy."""
    lengths = [2048, 8192, 16384]
    xis = [2.0, 4.0, 8.0]
    models = [0, 1, 0]
    noises, oracle, params, h0 = make_problem(q, lengths, 2048, xis, models, [21, 22, 23], q.F64)
    shifts = [0.5 / xi ** 2 for xi in xis]
    refs = []
    for k in range(3):
        on, po = oracle[k]
        h_ref, r_ref = O.lbfgs(h0[k], po, on, O.lbfgs_params(m=10, ftol=1e-5, max_iter=4000,
                                                             precondition=True,
                                                             precond_shift=shifts[k]), omp=True)
        assert r_ref.converged
        refs.append((h_ref, r_ref, O.objective(h_ref, po, on)))
    for dtype, tol in ((q.F64, 1e-10), (q.F32, 1e-6)):
        nz = noises if dtype == q.F64 else [n.convert(q.F32) for n in noises]
        b = q.Batch(lengths, nz, params)
        b.set_precond_shifts(shifts)
        h = dev(np.concatenate(h0))
        # FP32 line offsets resolve ~ulp(|offset|) ~ 4e-6 cells at 50 cells from the
        # anchor, which floors the FP32 gradient near 1e-4 x curvature scale:
        # FP32 runs stop at ftol 1e-4 and are checked by the energy they reach
        ftol = 1e-5 if dtype == q.F64 else 1e-4
        res = b.lbfgs(h, q.LbfgsParams(m=10, ftol=ftol, max_iter=4000, precondition=True),
                      dtype=dtype)
        hh = host(h)
        for k in range(3):
            on, po = oracle[k]
            h_ref, r_ref, e_ref = refs[k]
            assert res[k].converged, (dtype, k, res[k])
            if dtype == q.F64:
                assert abs(res[k].iterations - r_ref.iterations) <= max(5, r_ref.iterations // 5)
            e = O.objective(hh[b.offsets[k]:b.offsets[k + 1]], po, on)
            assert rel(e, e_ref) < tol, (dtype, k, e, e_ref)
