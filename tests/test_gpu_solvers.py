"""GPU parity of the solvers (HIP path through the C ABI) against the CPU
oracle: L-BFGS (host-driven FP64 path), Langevin steps (Philox normals) and
the Rosso-Krauth construction, plus the reference's own solver tests
restated on the GPU (tests/test_lbfgs.cpp, test_dynamics.cpp,
test_rosso_krauth.cpp)."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, rel, torch_cuda

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def q():
    torch_cuda()
    import qew_amd
    return qew_amd


# ---- L-BFGS ------------------------------------------------------------------

def test_lbfgs_golden_fixture(q):
    g = json.load(open(os.path.join(GOLD, "lbfgs_golden.json")))
    nx, ny = int(g["nx"]), int(g["ny"])
    field = np.array(g["field"]["data"]).reshape(nx, ny)
    gn = q.FilteredNoise.from_field(field, g["Lx"], g["Ly"])
    p = q.Params(g["Lx"], g["line_tension"], 0.0, q.Model.Linear)
    h = dev(np.array(g["h0"]["data"]))
    r = q.lbfgs_minimize(h, p, gn, q.LbfgsParams(ftol=1e-8, precondition=True))
    assert r.converged
    e = q.objective(h, p, gn)
    assert abs(e - g["e_min"]) < 1e-5 * abs(g["e_min"]) + 1e-6


@pytest.mark.parametrize("precondition", [False, True])
@pytest.mark.parametrize("model", [0, 1])
def test_lbfgs_matches_oracle(q, precondition, model):
    n, ny = 256, 64
    field, coef = O.filtered_noise(n, ny, 2.56, 1.0, 0.3, 0.2, 0.2, 5)
    on = O.OracleNoise.table(coef, 2.56, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 0.01, 1.0 / ny)
    lt = 50.0 if precondition else 1.0
    p_o = O.params(2.56, lt, 0.0, model)
    p = q.Params(2.56, lt, 0.0, q.Model(model))
    # ftol stays above the Armijo roundoff floor of each problem (below it
    # the sufficient-decrease test compares rounding noise and both sides
    # stall), so both sides converge; the minima then agree to the energy's
    # rounding and to ftol / (smallest curvature) in position
    ftol = 1e-8 if precondition else 1e-6
    lpo = O.lbfgs_params(m=10, ftol=ftol, max_iter=5000, precondition=precondition,
                         precond_shift=0.5 * 0.01 * 0.3 / 0.04 if precondition else 0.0)
    lp = q.LbfgsParams(m=10, ftol=ftol, max_iter=5000, precondition=precondition,
                       precond_shift=lpo.precond_shift)
    h_ref, r_ref = O.lbfgs(np.full(n, 0.5), p_o, on, lpo)
    h = dev(np.full(n, 0.5))
    r = q.lbfgs_minimize(h, p, gn, lp)
    assert r.converged and r_ref.converged
    assert abs(r.iterations - r_ref.iterations) <= max(2, r_ref.iterations // 10)
    e_ref = O.objective(h_ref, p_o, on)
    assert rel(q.objective(h, p, gn), e_ref) < 1e-11
    assert np.max(np.abs(host(h) - h_ref)) < 1e-5


def test_lbfgs_branches(q):
    well = q.HarmonicWell(2.0, 0.5)
    lin = lambda n: q.Params(float(n), 1.0, 0.0, q.Model.Linear)
    # relaxes to flat / harmonic minimum (tests/test_lbfgs.cpp:67-124)
    h = dev(0.5 + 0.3 * np.cos(4 * np.pi * np.arange(128) / 128))
    assert q.lbfgs_minimize(h, lin(128), q.ZeroNoise(), q.LbfgsParams(ftol=1e-9)).converged
    assert np.ptp(host(h)) < 1e-6
    for model in (q.Model.Linear, q.Model.Arclength):
        h = dev(np.zeros(96))
        p = q.Params(96.0, 1.0, 0.0, model)
        assert q.lbfgs_minimize(h, p, well, q.LbfgsParams(ftol=1e-9)).converged
        assert np.max(np.abs(host(h) - 0.5)) < 1e-6
    # non-convergence, curvature skip, no-history give-up (:126-188)
    r = q.lbfgs_minimize(dev(np.zeros(96)), lin(96), well, q.LbfgsParams(ftol=1e-12, max_iter=1))
    assert not r.converged and r.max_force >= 1e-12
    h = dev(np.zeros(16))
    r = q.lbfgs_minimize(h, lin(16), well, q.LbfgsParams(ftol=1e-8, max_iter=20000, curvature_eps=1e30))
    assert r.n_curvature_skips > 0 and r.converged
    r = q.lbfgs_minimize(dev(np.zeros(32)), lin(32), well, q.LbfgsParams(ftol=1e-9, max_ls=0))
    assert not r.converged and r.iterations == 0 and r.n_ls_restarts == 0


def test_lbfgs_restart_branch(q):
    # tests/test_lbfgs.cpp:200-228 on the device-generated table
    gn = q.FilteredNoise(64, 64, 3.2, 1.0, 2.0, 0.15, 0.15, 3)
    p = q.Params(3.2, 0.5, 0.0, q.Model.Linear)
    r = q.lbfgs_minimize(dev(np.full(64, 0.5)), p, gn, q.LbfgsParams(ftol=1e-6, max_iter=50000, max_ls=3))
    assert r.n_ls_restarts > 0
    r2 = q.lbfgs_minimize(dev(np.full(64, 0.5)), p, gn, q.LbfgsParams(ftol=1e-6, max_iter=50000))
    assert r2.converged and r2.n_ls_restarts == 0


# ---- Langevin ----------------------------------------------------------------

def test_langevin_steps_match_oracle(q):
    n = 1000
    field, coef = O.filtered_noise(n, 64, 10.0, 1.0, 0.5, 0.2, 0.2, 8)
    on = O.OracleNoise.table(coef, 10.0, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 0.01, 1.0 / 64)
    for model in (0, 1):
        p_o = O.params(10.0, 0.7, 0.2, model)
        p = q.Params(10.0, 0.7, 0.2, q.Model(model))
        h_ref = np.random.default_rng(3).uniform(0, 1, n)
        h = dev(h_ref)
        dyn = q.LangevinDynamics(n, 77)
        for s in range(20):
            O.langevin_step(h_ref, p_o, on, 0.003, 0.01, 77, s)
            dyn.step(h, p, gn, 0.003, 0.01)
        # normals use device log/sin/cos: agreement to a few ulp per step
        assert np.max(np.abs(host(h) - h_ref)) < 1e-13


def test_langevin_zero_diffusion_bitwise_and_seed_free(q):
    n = 48
    field, coef = O.filtered_noise(n, 32, 4.8, 1.0, 0.5, 0.3, 0.3, 2)
    on = O.OracleNoise.table(coef, 4.8, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 4.8 / n, 1.0 / 32)
    p_o = O.params(4.8, 1.0, 0.3, 0)
    p = q.Params(4.8, 1.0, 0.3, q.Model.Linear)
    h0 = np.random.default_rng(5).uniform(-0.5, 0.5, n)
    h1, h2, h_ref = dev(h0), dev(h0), h0.copy()
    d1, d2 = q.LangevinDynamics(n, 111), q.LangevinDynamics(n, 999)
    for s in range(50):
        d1.step(h1, p, gn, 0.05, 0.0)
        d2.step(h2, p, gn, 0.05, 0.0)
        O.langevin_step(h_ref, p_o, on, 0.05, 0.0, 0, s)
    np.testing.assert_array_equal(host(h1), host(h2))
    np.testing.assert_array_equal(host(h1), h_ref)


def test_langevin_force_shifts_equilibrium(q):
    # tests/test_dynamics.cpp:106-130
    k, y0, f = 4.0, 0.2, 0.6
    h = dev(np.zeros(32))
    dyn = q.LangevinDynamics(32, 0)
    p = q.Params(32.0, 1.0, f, q.Model.Linear)
    well = q.HarmonicWell(k, y0)
    for _ in range(5000):
        dyn.step(h, p, well, 0.05, 0.0)
    assert np.max(np.abs(host(h) - (y0 + f / k))) < 1e-6


def test_batch_langevin_matches_single(q):
    torch = torch_cuda()
    n = 512
    field, coef = O.filtered_noise(n, 64, 5.12, 1.0, 0.5, 0.2, 0.2, 4)
    gn = q.FilteredNoise.from_coefficients(coef, 0.01, 1.0 / 64)
    p = q.Params(5.12, 0.9, 0.1, q.Model.Arclength)
    h0 = np.random.default_rng(9).uniform(0, 1, n)
    b = q.Batch([n, n], [gn, gn], [p, p])
    hb = dev(np.concatenate([h0, h0]))
    steps = 7
    b.langevin(hb, [0.002] * steps, [0.01] * steps, seed=40)
    hs = dev(h0)
    dyn = q.LangevinDynamics(n, 40)
    for _ in range(steps):
        dyn.step(hs, p, gn, 0.002, 0.01)
    np.testing.assert_array_equal(host(hb)[:n], host(hs))
    assert np.max(np.abs(host(hb)[n:] - host(hs))) > 0  # line 1 uses seed + 1


@pytest.mark.parametrize("table,k,steps,tile", [
    ("f64", "16", 37, "block"), ("f32", "16", 37, "block"), ("f64", "8", 37, "block"),
    ("f64", "32", 37, "block"), ("f64", "8", 37, "wave"), ("f64", "16", 37, "wave"),
    ("f32", "8", 37, "wave"), ("f32", "16", 37, "wave"), ("f64", None, 100, None)])
def test_batch_langevin_multistep_tiles_bitwise(q, monkeypatch, table, k, steps, tile):
    # several steps per launch on workgroup or wave tiles with a halo (lines
    # shorter and longer than a tile, odd lengths, both models, a partial last
    # launch, steps without noise), for each launch plan and for the timed
    # choice (>= 96 steps): every line equals single steps of the same stream
    # bit for bit
    for name, val in (("QEW_LANGEVIN_STEPS_PER_LAUNCH", k), ("QEW_LANGEVIN_TILE", tile)):
        if val is None:
            monkeypatch.delenv(name, raising=False)
        else:
            monkeypatch.setenv(name, val)
    field, coef = O.filtered_noise(256, 64, 2.56, 1.0, 0.5, 0.2, 0.2, 6)
    gn = q.FilteredNoise.from_coefficients(coef, 0.01, 1.0 / 64)
    if table == "f32":
        gn = gn.convert(q.F32)
    lengths = [5, 100, 224, 1000, 2]
    params = [q.Params(2.56, 0.8, 0.05, q.Model(k % 2)) for k in range(len(lengths))]
    b = q.Batch(lengths, [gn] * len(lengths), params)
    rng = np.random.default_rng(12)
    h0 = [rng.uniform(0.3, 0.7, n) for n in lengths]
    drift = np.full(steps, 0.002)
    diff = np.where(np.arange(steps) % 5 == 3, 0.0, 0.01)
    hb = dev(np.concatenate(h0))
    b.langevin(hb, drift, diff, seed=21, step0=1000)
    off = np.concatenate([[0], np.cumsum(lengths)])
    for k, n in enumerate(lengths):
        hs = dev(h0[k])
        dyn = q.LangevinDynamics(n, 21 + k)
        dyn.steps = 1000
        for s in range(steps):
            dyn.step(hs, params[k], gn, drift[s], diff[s])
        np.testing.assert_array_equal(host(hb)[off[k]:off[k + 1]], host(hs), err_msg=f"line {k} n={n}")


@pytest.mark.parametrize("tile", ["block", "wave"])
def test_batch_langevin_fp32_multistep_matches_single_steps(q, monkeypatch, tile):
    # FP32 line state: a multi-step call equals the same steps issued one per
    # call (two lines over several tiles, one odd line shorter than a tile)
    monkeypatch.setenv("QEW_LANGEVIN_TILE", tile)
    monkeypatch.delenv("QEW_LANGEVIN_STEPS_PER_LAUNCH", raising=False)
    n, ny = 700, 256
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 5)
    g32 = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0, dtype=q.F32)
    p = [q.Params(float(n), 1.0, 0.05, q.Model.Linear), q.Params(float(n), 1.0, 0.05, q.Model.Arclength),
         q.Params(37.0, 1.0, 0.05, q.Model.Linear)]
    b = q.Batch([n, n, 37], [g32, g32, g32], p)
    anchor = dev(np.array([ny / 2.0, ny / 2.0 + 0.25, ny / 2.0 - 0.5]))
    o0 = np.random.default_rng(2).normal(0, 1, 2 * n + 37).astype(np.float32)
    steps = 40
    drift = np.full(steps, 0.01)
    diff = np.full(steps, np.sqrt(2 * 0.01 * 0.05))
    om = dev(o0)
    b.langevin(om, drift, diff, seed=3, dtype=q.F32, anchor=anchor)
    os_ = dev(o0)
    for s in range(steps):
        b.langevin(os_, drift[s:s + 1], diff[s:s + 1], seed=3, step0=s, dtype=q.F32, anchor=anchor)
    np.testing.assert_array_equal(host(om), host(os_))


# ---- Rosso-Krauth --------------------------------------------------------------

def test_rk_block_matches_oracle_bitwise(q):
    for model, seed, amp, lt, f in [(0, 7, 0.5, 1.0, 0.0), (1, 29, 0.6, 1.2, 0.1), (1, 9, 0.5, 1.5, 0.3)]:
        field, coef = O.filtered_noise(64, 64, 3.2, 1.0, amp, 0.2, 0.2, seed)
        on = O.OracleNoise.table(coef, 3.2, 1.0)
        gn = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / 64)
        rpo = O.rk_params(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-11, sweep_tol=1e-9, runaway=3.0)
        rp = q.RkParams(0.5 / 64, 1.0, 1e-11, 1e-9, 3.0)
        h_ref = np.full(64, 0.5)
        r_ref = O.rk_block(h_ref, O.params(3.2, lt, f, model), on, rpo)
        h = dev(np.full(64, 0.5))
        r = q.rk_block(h, q.Params(3.2, lt, f, q.Model(model)), gn, rp)
        assert (r.blocked, r.runaway, r.sweeps) == (bool(r_ref.blocked), bool(r_ref.runaway), r_ref.sweeps)
        np.testing.assert_array_equal(host(h), h_ref)
        assert abs(r.mean_h - r_ref.mean_h) < 1e-12
        assert r.n_velocity_evals == r_ref.n_velocity_evals > 0


def test_rk_washboard_critical_force(q):
    # tests/test_rosso_krauth.cpp:49-89 (10 bisection steps)
    wash = q.SinusoidNoise(0.1, 2 * math.pi)
    rp = q.RkParams(1 / 200, 1.0, 1e-12, 1e-11, 3.0, 100000)
    fc = 0.2 * math.pi

    def blocks(f):
        return q.rk_block(dev(np.zeros(16)), q.Params(16.0, 1.0, f, q.Model.Linear), wash, rp).blocked

    lo, hi = 0.0, 2 * fc
    assert blocks(lo) and not blocks(hi)
    for _ in range(10):
        m = 0.5 * (lo + hi)
        lo, hi = (m, hi) if blocks(m) else (lo, m)
    assert abs(0.5 * (lo + hi) - fc) < 0.02 * fc
    r = q.rk_block(dev(np.zeros(16)), q.Params(16.0, 1.0, 2 * fc, q.Model.Linear), wash, rp)
    assert r.runaway and not r.blocked and r.mean_h > 3.0


def test_fc_bisection_to_1e10_matches_oracle(q):
    """The depinning driver's force ramp and f_c bisection (the facade's
    depinning_ramp: ramp f = k f_step carrying h over, then bisection of
    [last blocked, first unblocked] warm-started from the highest blocked
    configuration) run to 1e-10 relative width with the GPU rk_block and
    with the oracle's: identical decisions on every probe, identical bracket
    bits and critical configuration. Probes are capped at 2000 sweeps (a
    capped probe counts as not blocked, as in the driver), which defines the
    same critical force on both sides and keeps the run to seconds."""
    n = 32
    field, coef = O.filtered_noise(n, 64, n * 0.05, 1.0, 0.5, 0.2, 0.2, 7)
    on = O.OracleNoise.table(coef, n * 0.05, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / 64)
    kw = dict(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-11, sweep_tol=1e-9, runaway=1.0,
              max_sweeps=2000)
    rpo = O.rk_params(**kw)
    rp = q.RkParams(kw["dstep"], kw["max_advance"], kw["root_tol"], kw["sweep_tol"], kw["runaway"],
                    kw["max_sweeps"])

    def gpu(h, f):
        hd = dev(h)
        r = q.rk_block(hd, q.Params(n * 0.05, 1.0, f, q.Model.Linear), gn, rp)
        return bool(r.blocked), host(hd)

    def oracle(h, f):
        hh = h.copy()
        r = O.rk_block(hh, O.params(n * 0.05, 1.0, f, 0), on, rpo)
        return bool(r.blocked), hh

    def depin(rk, f_step=0.1, rel_tol=1e-10):
        h, f_lo, f_hi, h_lo, trace = np.full(n, 0.5), None, None, None, []
        for k in range(400):
            f = k * f_step
            blocked, h = rk(h, f)
            trace.append(blocked)
            if not blocked:
                f_hi = f
                break
            f_lo, h_lo = f, h.copy()
        lo, hi = f_lo, f_hi
        while hi - lo > rel_tol * max(abs(hi), abs(lo)):
            mid = 0.5 * (lo + hi)
            if not (lo < mid < hi):
                break
            blocked, hm = rk(h_lo, mid)
            trace.append(blocked)
            if blocked:
                lo, h_lo = mid, hm
            else:
                hi = mid
        return lo, hi, h_lo, trace

    lo_g, hi_g, hc_g, tr_g = depin(gpu)
    lo_o, hi_o, hc_o, tr_o = depin(oracle)
    assert tr_g == tr_o and len(tr_g) > 60
    assert (lo_g, hi_g) == (lo_o, hi_o)
    assert 0 < hi_g - lo_g <= 1e-10 * hi_g
    np.testing.assert_array_equal(hc_g, hc_o)


@pytest.mark.parametrize("path", ["line", "launch", "hybrid", "line/7", "hybrid/5"])
@pytest.mark.parametrize("n,lt,f,max_sweeps", [
    (16384, 0.005, 0.1, 3000),   # largest line the LDS path holds; runs to blocked
    (16384, 0.02, 0.2, 3000),
    (32768, 0.02, 0.2, 300),     # beyond LDS (per-colour launches), stops at max_sweeps
])
def test_rk_long_lines_match_oracle_bitwise(q, monkeypatch, path, n, lt, f, max_sweeps):
    # hybrid: per-colour launches for the first poll's sweeps, then the LDS
    # kernel continues the same call from the launches' state and flags;
    # "path/k": the LDS kernel stops after k sweeps per launch and the next
    # launch continues from the saved heights, moved flags and counters
    path, _, per_launch = path.partition("/")
    monkeypatch.setenv("QEW_RK_KERNEL", path)
    if per_launch:
        monkeypatch.setenv("QEW_RK_LINE_SWEEPS", per_launch)
    else:
        monkeypatch.delenv("QEW_RK_LINE_SWEEPS", raising=False)
    # 16 x 16384 and 64 x 32768 sites per column of a 256 x 64 table: many
    # moving sites per workgroup in the first sweeps, few near the end, so
    # the lane-group sizes of the cooperative march/bisection all occur
    field, coef = O.filtered_noise(256, 64, 256.0, 1.0, 0.5, 4.0, 0.1, 3)
    on = O.OracleNoise.table(coef, 256.0, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0 / 64)
    rpo = O.rk_params(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-10, sweep_tol=1e-9,
                      runaway=3.0, max_sweeps=max_sweeps)
    rp = q.RkParams(0.5 / 64, 1.0, 1e-10, 1e-9, 3.0, max_sweeps)
    h_ref = np.full(n, 0.5)
    r_ref = O.rk_block(h_ref, O.params(256.0, lt, f, 0), on, rpo, omp=True)
    h = dev(np.full(n, 0.5))
    r = q.rk_block(h, q.Params(256.0, lt, f, q.Model.Linear), gn, rp)
    assert (r.blocked, r.runaway, r.sweeps) == (bool(r_ref.blocked), bool(r_ref.runaway), r_ref.sweeps)
    assert r.n_velocity_evals == r_ref.n_velocity_evals
    # the device skips unchanged sites: it evaluates fewer velocities than the
    # reference algorithm, but at least one per site in the first sweep
    assert n <= r.n_velocity_evals_done
    np.testing.assert_array_equal(host(h), h_ref)
    assert abs(r.mean_h - r_ref.mean_h) < 1e-12


def test_rk_odd_length_rejected(q):
    with pytest.raises(ValueError):
        q.rk_block(dev(np.zeros(17)), q.Params(17.0, 1.0, 0.0, q.Model.Linear), q.SinusoidNoise(0.1, 6.28),
                   q.RkParams(0.01, 1.0, 1e-10, 1e-9, 3.0))


def test_rk_site_velocity_is_minus_gradient(q):
    torch = torch_cuda()
    field, coef = O.filtered_noise(64, 64, 3.2, 1.0, 0.7, 0.2, 0.2, 13)
    gn = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / 64)
    for model in (q.Model.Linear, q.Model.Arclength):
        p = q.Params(3.2, 1.3, 0.4, model)
        h = np.random.default_rng(21).uniform(0.1, 0.9, 64)
        g = host(q.gradient(dev(h), p, gn))
        idx = torch.arange(64, dtype=torch.int32, device="cuda:0")
        v = q.rk_site_velocity(idx, 64, p, gn, dev(h), dev(np.roll(h, 1)), dev(np.roll(h, -1)))
        assert np.max(np.abs(host(v) + g)) < 1e-12


@pytest.mark.parametrize("path", ["line", "launch", "hybrid", None])
def test_batch_rk_matches_single_lines(q, monkeypatch, path):
    # the batch runs on the chosen path (None: the path timed on the batch's
    # first call), the single lines on the default one
    field, coef = O.filtered_noise(64, 64, 3.2, 1.0, 0.5, 0.2, 0.2, 7)
    gn = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / 64)
    rp = q.RkParams(0.5 / 64, 1.0, 1e-11, 1e-9, 3.0)
    forces = [0.0, 0.05, 0.1, 5.0, 0.07]
    lengths = [64, 128, 32, 64, 1026]
    params = [q.Params(3.2, 1.0, f, q.Model.Linear) for f in forces]
    b = q.Batch(lengths, [gn] * len(lengths), params)
    hb = dev(np.full(sum(lengths), 0.5))
    if path is None:
        monkeypatch.delenv("QEW_RK_KERNEL", raising=False)
    else:
        monkeypatch.setenv("QEW_RK_KERNEL", path)
    res = b.rk_block(hb, rp)
    monkeypatch.delenv("QEW_RK_KERNEL", raising=False)
    off = np.concatenate([[0], np.cumsum(lengths)])
    for k, f in enumerate(forces):
        h = dev(np.full(lengths[k], 0.5))
        r = q.rk_block(h, params[k], gn, rp)
        assert (res[k].blocked, res[k].runaway, res[k].sweeps) == (r.blocked, r.runaway, r.sweeps)
        assert res[k].n_velocity_evals == r.n_velocity_evals
        np.testing.assert_array_equal(host(hb)[off[k]:off[k + 1]], host(h))


def test_rk_fp32_table_bitwise_vs_oracle_on_rounded_table(q):
    # FP32 storage, FP64 arithmetic: identical to the FP64 oracle run on the
    # FP32-rounded coefficients
    field, coef = O.filtered_noise(64, 64, 3.2, 1.0, 0.6, 0.2, 0.2, 29)
    coef32 = coef.astype(np.float32).astype(np.float64)
    on = O.OracleNoise.table(coef32, 3.2, 1.0)
    gn = q.FilteredNoise.from_coefficients(coef, 0.05, 1.0 / 64, dtype=q.F32)
    rpo = O.rk_params(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-11, sweep_tol=1e-9, runaway=3.0)
    h_ref = np.full(64, 0.5)
    r_ref = O.rk_block(h_ref, O.params(3.2, 1.2, 0.1, 1), on, rpo)
    h = dev(np.full(64, 0.5))
    r = q.rk_block(h, q.Params(3.2, 1.2, 0.1, q.Model.Arclength), gn, q.RkParams(0.5 / 64, 1.0, 1e-11, 1e-9, 3.0))
    assert (r.blocked, r.sweeps) == (bool(r_ref.blocked), r_ref.sweeps)
    np.testing.assert_array_equal(host(h), h_ref)


def test_batch_langevin_fp32_vs_fp64_statistics(q):
    """FP32 line state (offsets from an anchor) + FP32 table against FP64 on
    the same table values and the same Philox normals: mean velocity (mean
    displacement over the elapsed time) and roughness W^2 after 2000 steps
    agree to 1e-5 relative, the Langevin configuration's tolerance."""
    torch = torch_cuda()
    n, ny = 4096, 1024
    field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, 3)
    g64 = q.FilteredNoise.from_coefficients(coef.astype(np.float32).astype(np.float64), 1.0, 1.0)
    g32 = g64.convert(q.F32)
    p = q.Params(float(n), 1.0, 0.05, q.Model.Linear)
    steps = 2000
    drift = np.full(steps, 0.01)
    diff = np.full(steps, np.sqrt(2 * 0.01 * 0.05))
    h0 = np.full(n, ny / 2.0)
    b64 = q.Batch([n], [g64], [p])
    b32 = q.Batch([n], [g32], [p])
    h64 = dev(h0)
    b64.langevin(h64, drift, diff, seed=11)
    anchor = dev(np.array([ny / 2.0]))
    o32 = dev(np.zeros(n, dtype=np.float32))
    b32.langevin(o32, drift, diff, seed=11, dtype=q.F32, anchor=anchor)
    a = host(h64)
    b = ny / 2.0 + host(o32).astype(np.float64)
    t = steps * 0.01
    va, vb = (a.mean() - ny / 2.0) / t, (b.mean() - ny / 2.0) / t
    assert rel(vb, va) < 1e-5, (va, vb)
    w2a, w2b = a.var(), b.var()
    assert rel(w2b, w2a) < 1e-5, (w2a, w2b)
