"""Midpoint-contracted FP32 tables (QEW_F32_MID: storage only, FP32 line state).

T(ix, iy) = sum_a w_a c(ix - 1 + a mod nx, iy), with w the cubic B-spline
weights at x-fraction 1/2, (1, 23, 23, 1)/48, contracted in FP64 and rounded
to FP32 once. On a line whose site spacing equals the table's x spacing and
whose length is even, every edge midpoint lies at x-fraction 1/2: interior
edges at i + 1/2 and the wrap edge at (n - 1)/2 (the reference's wrap-edge
convention, src/qew_model.h:79). So V and dV/dy at an edge are the y-spline of
4 values of T instead of the 4 x 4 coefficient window. The tests pin the
contraction against numpy, the evaluation and dynamics against the FP64 oracle
(the stated FP32 tolerances), and the plans of the batched dynamics against
each other bit for bit; every other entry point and every other line geometry
is rejected.
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


def mid_fold(coef):
    """numpy contraction of row-major coef[ix, iy] with the B-spline weights
    at t = 1/2 (the textbook cubic B-spline basis), FP64, a = 0..3 in order."""
    t = 0.5
    w = [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6,
         (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]
    acc = np.zeros_like(coef)
    for a in range(4):
        acc = acc + w[a] * np.roll(coef, 1 - a, axis=0)  # coef[(ix - 1 + a) mod nx]
    return acc


def test_mid_weights_are_the_midpoint_bspline():
    """The weights the contraction uses: (1, 23, 23, 1) / 48, summing to 1."""
    t = 0.5
    w = np.array([(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6,
                  (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6])
    np.testing.assert_allclose(w * 48, [1, 23, 23, 1], rtol=0, atol=1e-13)
    assert abs(w.sum() - 1) < 1e-15


def test_mid_table_values(q):
    for nx, ny, seed in [(64, 48, 3), (67, 33, 5)]:
        field, coef = O.filtered_noise(nx, ny, float(nx), float(ny), 1.0, 4.0, 4.0, seed)
        g = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
        m = g.convert(q.F32_MID)
        assert m.info()["dtype"] == q.F32_MID
        ref = mid_fold(coef).astype(np.float32)
        np.testing.assert_array_max_ulp(m.coefficients().astype(np.float32), ref, maxulp=1)
        # from an FP32 table: the contraction of its rounded coefficients
        m32 = g.convert(q.F32).convert(q.F32_MID)
        ref32 = mid_fold(coef.astype(np.float32).astype(np.float64)).astype(np.float32)
        np.testing.assert_array_max_ulp(m32.coefficients().astype(np.float32), ref32, maxulp=1)


def test_mid_batch_eval_matches_oracle(q):
    """Ragged batch on one table (spacing 1 = table dx): the full-length line
    (wrap edge in the table's middle), a half-length line, and an arclength
    line with a force; gradient and energy against the FP64 oracle within the
    FP32 path's stated tolerance (2e-4 of max|grad|, 1e-5 relative), and no
    worse than the FP32 table's own error by more than a factor 2."""
    n, ny = 4096, 256
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 21)
    on = O.OracleNoise.table(coef, float(n), float(ny))
    g64 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
    gmid, g32 = g64.convert(q.F32_MID), g64.convert(q.F32)
    lengths = [n, n // 2, n]
    models = [q.Model.Linear, q.Model.Arclength, q.Model.Arclength]
    forces = [0.0, 0.0, 0.3]
    params = [q.Params(float(L), 1.0, f, m) for L, f, m in zip(lengths, forces, models)]
    rng = np.random.default_rng(8)
    hs = [ny / 2 + 0.37 * k + np.cumsum(rng.standard_normal(L)) for k, L in enumerate(lengths)]
    anchor = np.array([ny / 2 + 0.37 * k for k in range(3)])
    off = dev(np.concatenate([(h - a).astype(np.float32) for h, a in zip(hs, anchor)]))
    em, gm = q.Batch(lengths, [gmid] * 3, params).eval(off, dtype=q.F32, anchor=dev(anchor))
    e3, g3 = q.Batch(lengths, [g32] * 3, params).eval(off, dtype=q.F32, anchor=dev(anchor))
    gm, g3, em, e3 = host(gm), host(g3), host(em), host(e3)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    for k, (L, f, m) in enumerate(zip(lengths, forces, models)):
        # the oracle on the FP32 offsets the device sees
        h = anchor[k] + off[offs[k]:offs[k + 1]].cpu().numpy().astype(np.float64)
        po = O.params(float(L), 1.0, f, int(m))
        g_ref, e_ref = O.gradient(h, po, on), O.objective(h, po, on)
        gmax = np.max(np.abs(g_ref))
        err_mid = np.max(np.abs(gm[offs[k]:offs[k + 1]] - g_ref))
        err_32 = np.max(np.abs(g3[offs[k]:offs[k + 1]] - g_ref))
        assert err_mid < 2e-4 * gmax, (k, err_mid / gmax)
        assert err_mid <= 2 * err_32 + 1e-7 * gmax, (k, err_mid, err_32)
        assert rel(em[k], e_ref) < 1e-5, (k, rel(em[k], e_ref))


def test_mid_batch_langevin_plans_bitwise_and_close_to_fp64(q):
    """Every launch plan gives the same bits as single steps on a midpoint
    table, and the trajectory stays within the FP32 state's tolerance of the
    FP64 dynamics on the FP64 table at the same Philox stream (heights within
    1e-4, the squared width W^2 within 1e-4 relative after 64 steps)."""
    import os
    n, ny = 2048, 256
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 13)
    g64 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
    gmid = g64.convert(q.F32_MID)
    p = q.Params(float(n), 1.0, 0.05, q.Model.Linear)
    steps = 64
    drift = np.full(steps, 0.01)
    diff = np.full(steps, 0.05)
    h0 = np.random.default_rng(3).normal(0, 1, n)
    anchor = dev(np.array([ny / 2.0]))
    ref = dev(h0.astype(np.float32))
    for s in range(steps):  # one step per call: the single-step kernel
        q.Batch([n], [gmid], [p]).langevin(ref, drift[s:s + 1], diff[s:s + 1], seed=4, step0=s,
                                           dtype=q.F32, anchor=anchor)
    for tile, k in [("block", "8"), ("block", "16"), ("block", "32"), ("wave", "8"),
                    ("wave", "16")]:
        os.environ["QEW_LANGEVIN_TILE"], os.environ["QEW_LANGEVIN_STEPS_PER_LAUNCH"] = tile, k
        try:
            h = dev(h0.astype(np.float32))
            q.Batch([n], [gmid], [p]).langevin(h, drift, diff, seed=4, dtype=q.F32, anchor=anchor)
        finally:
            del os.environ["QEW_LANGEVIN_TILE"], os.environ["QEW_LANGEVIN_STEPS_PER_LAUNCH"]
        np.testing.assert_array_equal(host(h), host(ref), err_msg=f"{tile} {k}")
    h64 = dev(ny / 2.0 + h0)
    q.Batch([n], [g64], [p]).langevin(h64, drift, diff, seed=4)
    a = host(h64) - ny / 2.0
    b = host(ref).astype(np.float64)
    assert np.max(np.abs(b - a)) < 1e-4
    assert rel(b.var(), a.var()) < 1e-4


def test_mid_batch_lbfgs_energy_error(q):
    """The batched FP32 minimisation on midpoint tables (the FP32-table test's
    problem, test_gpu_batch_lbfgs.py): every minimum within 1e-6 relative
    energy of the FP64 oracle's, evaluated by the oracle."""
    LT = 4.0 ** 1.5
    lengths = [1024, 2048, 512, 1024, 4096, 256]
    xis = [2.0, 4.0, 8.0, 4.0, 8.0, 2.0]
    models = [0, 1, 0, 1, 1, 0]
    ny = 256
    noises, oracle, params, h0 = [], [], [], []
    for n, xi, model, seed in zip(lengths, xis, models, [5, 6, 7, 8, 9, 10]):
        field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, xi, xi, seed)
        oracle.append((O.OracleNoise.table(coef, float(n), float(ny)), O.params(float(n), LT, 0.0, model)))
        noises.append(q.FilteredNoise.from_coefficients(coef, 1.0, 1.0).convert(q.F32_MID))
        params.append(q.Params(float(n), LT, 0.0, q.Model(model)))
        h0.append(np.full(n, ny / 2.0))
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
        e, e_ref = O.objective(line, po, on), O.objective(h_ref, po, on)
        assert rel(e, e_ref) < 1e-6, (k, res[k], e, e_ref)
    # FP64 line state does not read midpoint tables
    with pytest.raises(q.QewInvalidArgument):
        b.lbfgs(dev(np.concatenate(h0)), q.LbfgsParams(m=10, precondition=True), dtype=q.F64)


@pytest.mark.parametrize("precondition", [True, False])
def test_f64_mid_batch_lbfgs_matches_oracle(q, precondition):
    """FP64 line state on FP64 midpoint tables against the oracle's L-BFGS on
    its MIDTABLE sampler with the same values: minima within 1e-11 relative
    energy (the FP64 batch's bar against the oracle: tree against sequential
    sums), all converged."""
    LT = 4.0 ** 1.5
    lengths = [256, 64, 512, 128]
    xis = [2.0, 4.0, 8.0, 4.0]
    models = [0, 1, 0, 1]
    ny = 128
    noises, oracle, params, h0 = [], [], [], []
    for n, xi, model, seed in zip(lengths, xis, models, [1, 2, 3, 4]):
        field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, xi, xi, seed)
        m64 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0).convert(q.F64_MID)
        assert m64.info()["dtype"] == q.F64_MID
        np.testing.assert_allclose(m64.coefficients(), mid_fold(coef), rtol=0,
                                   atol=1e-14 * np.max(np.abs(coef)))
        oracle.append((O.OracleNoise.mid_table(m64.coefficients(), float(n), float(ny)),
                       O.params(float(n), LT, 0.0, model)))
        noises.append(m64)
        params.append(q.Params(float(n), LT, 0.0, q.Model(model)))
        h0.append(np.full(n, ny / 2.0))
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
        e, e_ref = O.objective(line, po, on), O.objective(h_ref, po, on)
        assert rel(e, e_ref) < 1e-11, (k, e, e_ref)
    # read by no other entry point
    with pytest.raises(q.QewInvalidArgument):
        b.eval(dev(np.concatenate(h0)))
    with pytest.raises(q.QewInvalidArgument):
        b.lbfgs(dev(np.concatenate(h0)), q.LbfgsParams(m=4), dtype=q.F32)


@pytest.mark.parametrize("path", ["line", "launch", "hybrid"])
def test_mid_batch_rk_bitwise_vs_oracle(q, monkeypatch, path):
    """Rosso-Krauth sweeps on a midpoint table (FP64 arithmetic on the FP32
    values, as on FP32 tables) are bit for bit the oracle's on the same
    values (its MIDTABLE sampler, pinned against the full bicubic on the CPU):
    heights, blocked/runaway flags, sweeps and velocity-evaluation counts, on
    both device paths, for lines that block and a line that runs away."""
    nx, ny = 64, 64
    field, coef = O.filtered_noise(nx, ny, 3.2, 1.0, 0.6, 0.2, 0.2, 29)
    gmid = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / ny).convert(q.F32_MID)
    on = O.OracleNoise.mid_table(gmid.coefficients(), 3.2, 1.0)
    forces = [0.0, 0.1, 5.0, 20.0]  # the last runs away
    models = [q.Model.Linear, q.Model.Arclength, q.Model.Linear, q.Model.Arclength]
    params = [q.Params(3.2, 1.2, f, m) for f, m in zip(forces, models)]
    rp = q.RkParams(0.5 / 64, 1.0, 1e-11, 1e-9, 3.0)
    rpo = O.rk_params(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-11, sweep_tol=1e-9, runaway=3.0)
    b = q.Batch([nx] * len(forces), [gmid] * len(forces), params)
    hb = dev(np.full(nx * len(forces), 0.5))
    monkeypatch.setenv("QEW_RK_KERNEL", path)
    res = b.rk_block(hb, rp)
    monkeypatch.delenv("QEW_RK_KERNEL", raising=False)
    hh = host(hb)
    for k, (f, m) in enumerate(zip(forces, models)):
        h_ref = np.full(nx, 0.5)
        r_ref = O.rk_block(h_ref, O.params(3.2, 1.2, f, int(m)), on, rpo)
        assert (res[k].blocked, res[k].runaway, res[k].sweeps) == \
            (bool(r_ref.blocked), bool(r_ref.runaway), r_ref.sweeps), k
        assert res[k].n_velocity_evals == r_ref.n_velocity_evals, k
        np.testing.assert_array_equal(hh[k * nx:(k + 1) * nx], h_ref, err_msg=str(k))
    assert any(r.blocked for r in res) and any(r.runaway for r in res)


def test_mid_tables_rejected_elsewhere_and_for_other_geometries(q):
    n, ny = 64, 32
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 11)
    gmid = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0).convert(q.F32_MID)
    p = q.Params(float(n), 1.0, 0.0, q.Model.Linear)
    h = dev(np.full(n, ny / 2.0))
    rk = q.RkParams(0.5, 1.0, 1e-10, 1e-9, 2.0)
    for call in (lambda: q.gradient(h, p, gmid), lambda: q.objective(h, p, gmid),
                 lambda: q.lbfgs_minimize(h, p, gmid), lambda: q.rk_block(h, p, gmid, rk),
                 lambda: q.LangevinDynamics(n, 1).step(h, p, gmid, 0.01, 0.0),
                 lambda: gmid.sample(dev(np.array([1.0])), dev(np.array([1.0]))),
                 lambda: gmid.convert(q.F32), lambda: gmid.convert(q.F32_MID)):
        with pytest.raises(q.QewInvalidArgument):
            call()
    b = q.Batch([n], [gmid], [p])
    for call in (lambda: b.eval(h), lambda: b.lbfgs(h, q.LbfgsParams(m=4, precondition=False)),
                 lambda: b.langevin(h, [0.01], [0.0], seed=1)):
        with pytest.raises(q.QewInvalidArgument):
            call()
    off = dev(np.zeros(n, dtype=np.float32))
    anchor = dev(np.array([ny / 2.0]))
    b.eval(off, dtype=q.F32, anchor=anchor)  # the accepted geometry
    # site spacing != table dx, and an odd length (wrap edge off the midpoint)
    for L, phys in [(n, 2.0 * n), (n - 1, float(n - 1))]:
        bb = q.Batch([L], [gmid], [q.Params(phys, 1.0, 0.0, q.Model.Linear)])
        o = dev(np.zeros(L, dtype=np.float32))
        with pytest.raises(q.QewInvalidArgument):
            bb.eval(o, dtype=q.F32, anchor=anchor)
        with pytest.raises(q.QewInvalidArgument):
            bb.langevin(o, [0.01], [0.0], seed=1, dtype=q.F32, anchor=anchor)
        with pytest.raises(q.QewInvalidArgument):
            bb.lbfgs(dev(np.full(L, ny / 2.0)), q.LbfgsParams(m=4, precondition=False), dtype=q.F32)
        if L % 2 == 0:  # (odd lengths are rejected by the red-black sweep anyway)
            with pytest.raises(q.QewInvalidArgument):
                bb.rk_block(dev(np.full(L, ny / 2.0)), rk)


def test_f64_mid_tables_rejected_elsewhere_and_for_other_geometries(q):
    """QEW_F64_MID tables are read only by the batched FP64 L-BFGS, and only
    for lines of their geometry (even n, site spacing L / n equal to the
    table's dx): a wrong geometry would otherwise read the contracted column
    of the wrong cell without any fault, so every other use is rejected."""
    n, ny = 64, 32
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 12)
    g64 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
    gm = g64.convert(q.F64_MID)
    p = q.Params(float(n), 1.0, 0.0, q.Model.Linear)
    h = dev(np.full(n, ny / 2.0))
    rk = q.RkParams(0.5, 1.0, 1e-10, 1e-9, 2.0)
    fp = q.FireParams(dt_max=0.1)
    for call in (lambda: q.gradient(h, p, gm), lambda: q.objective(h, p, gm),
                 lambda: q.evaluate(h, p, gm), lambda: q.lbfgs_minimize(h, p, gm),
                 lambda: q.rk_block(h, p, gm, rk), lambda: q.fire_minimize(h, p, gm, fp),
                 lambda: q.LangevinDynamics(n, 1).step(h, p, gm, 0.01, 0.0),
                 lambda: gm.sample(dev(np.array([1.0])), dev(np.array([1.0]))),
                 lambda: gm.convert(q.F32), lambda: gm.convert(q.F64_MID)):
        with pytest.raises(q.QewInvalidArgument):
            call()
    b = q.Batch([n], [gm], [p])
    off = dev(np.zeros(n, dtype=np.float32))
    anchor = dev(np.array([ny / 2.0]))
    for call in (lambda: b.eval(h), lambda: b.eval(off, dtype=q.F32, anchor=anchor),
                 lambda: b.langevin(h, [0.01], [0.0], seed=1),
                 lambda: b.langevin(off, [0.01], [0.0], seed=1, dtype=q.F32, anchor=anchor),
                 lambda: b.lbfgs(h, q.LbfgsParams(m=4, precondition=False), dtype=q.F32),
                 lambda: b.rk_block(h, rk), lambda: b.fire(h, fp)):
        with pytest.raises(q.QewInvalidArgument):
            call()
    # the accepted geometry
    (r,) = b.lbfgs(dev(np.full(n, ny / 2.0)), q.LbfgsParams(m=4, ftol=1e-6), dtype=q.F64)
    assert r.converged
    # site spacing != table dx, and an odd length without the preconditioner
    for L, phys in [(n, 2.0 * n), (n - 1, float(n - 1))]:
        bb = q.Batch([L], [gm], [q.Params(phys, 1.0, 0.0, q.Model.Linear)])
        with pytest.raises(q.QewInvalidArgument):
            bb.lbfgs(dev(np.full(L, ny / 2.0)), q.LbfgsParams(m=4, precondition=False),
                     dtype=q.F64)
    # mixed FP64 / FP64_MID batches, in both orders
    for noises in ([g64, gm], [gm, g64]):
        bb = q.Batch([n, n], noises, [p, p])
        with pytest.raises(q.QewInvalidArgument):
            bb.lbfgs(dev(np.full(2 * n, ny / 2.0)), q.LbfgsParams(m=4, precondition=False),
                     dtype=q.F64)
