"""Pins the oracle's objective/gradient, L-BFGS, Philox/Langevin and
Rosso-Krauth against closed forms, finite differences, the lbfgs_golden
fixture and published known-answer vectors (reference tests/test_qew_model.cpp,
tests/test_lbfgs.cpp, tests/test_dynamics.cpp, tests/test_rosso_krauth.cpp).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ZERO = O.OracleNoise(O.ZERO)


def harmonic(k, y0):
    return O.OracleNoise(O.HARMONIC, p=(k, y0, 0, 0))


def washboard(A=0.1, k=2 * math.pi):
    return O.OracleNoise(O.WASHBOARD, p=(A, k, 0, 0))


# ---- objective / gradient -------------------------------------------------

@pytest.mark.parametrize("model", [O.LINEAR, O.ARCLENGTH])
def test_gradient_matches_fd_analytic_noise(model):
    # reference tests/test_qew_model.cpp:44-100
    n = 64
    p = O.params(n, 1.0, 0.37, model)
    nz = O.OracleNoise(O.SINCOS, p=(0.3, 0.7, 1.3, 0))
    h = np.random.default_rng(12345).uniform(0, 1, n)
    g = O.gradient(h, p, nz)
    eps = 1e-6
    fd = np.empty(n)
    for j in range(n):
        hp, hm = h.copy(), h.copy()
        hp[j] += eps
        hm[j] -= eps
        fd[j] = (O.objective(hp, p, nz) - O.objective(hm, p, nz)) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-6


@pytest.mark.parametrize("model", [O.LINEAR, O.ARCLENGTH])
def test_gradient_matches_fd_table_noise(model):
    # reference tests/test_filtered_noise.cpp:187-230
    _, coef = O.filtered_noise(64, 64, 3.2, 1.0, 0.5, 0.3, 0.3, 11)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    p = O.params(3.2, 0.5, 0.1, model)
    h = np.random.default_rng(7).uniform(0.2, 0.8, 64)
    g = O.gradient(h, p, nz)
    eps = 1e-6
    for j in range(64):
        hp, hm = h.copy(), h.copy()
        hp[j] += eps
        hm[j] -= eps
        fd = (O.objective(hp, p, nz) - O.objective(hm, p, nz)) / (2 * eps)
        assert abs(fd - g[j]) < 1e-5


def test_objective_flat_line_constant_potential():
    # reference tests/test_qew_model.cpp:110-132
    for model in (O.LINEAR, O.ARCLENGTH):
        p = O.params(3.0, 2.3, 0.4, model)
        e = O.objective(np.full(50, 0.7), p, O.OracleNoise(O.CONST, p=(1.25, 0, 0, 0)))
        assert abs(e - 3.0 * (1.25 - 0.4 * 0.7)) < 1e-12


def test_objective_sinusoid_linear_closed_form():
    # reference tests/test_qew_model.cpp:140-163
    n, lt, a = 32, 1.7, 0.9
    h = a * np.sin(2 * np.pi * np.arange(n) / n)
    e = O.objective(h, O.params(n, lt, 0.0, O.LINEAR), O.OracleNoise(O.CONST, p=(0, 0, 0, 0)))
    assert abs(e - (lt * a * a * n / 2) * (1 - np.cos(2 * np.pi / n))) < 1e-10


def test_objective_arclength_closed_form():
    # sum dx * lt * (sqrt(1 + s^2) - 1) for a sawtooth with |s| constant
    n, L, lt, amp = 40, 10.0, 1.3, 0.1
    h = amp * (np.arange(n) % 2)
    e = O.objective(h, O.params(L, lt, 0.0, O.ARCLENGTH), ZERO)
    s = amp / (L / n)
    assert abs(e - L * lt * (np.sqrt(1 + s * s) - 1)) < 1e-12


def test_objective_trial_equals_objective_of_materialised_point():
    rng = np.random.default_rng(3)
    _, coef = O.filtered_noise(32, 32, 3.2, 1.0, 0.5, 0.3, 0.3, 2)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    p = O.params(3.2, 0.7, 0.2, O.ARCLENGTH)
    x0, d = rng.uniform(0, 1, 32), rng.standard_normal(32)
    assert O.objective_trial(x0, d, 0.37, p, nz) == O.objective(x0 + 0.37 * d, p, nz)


# ---- preconditioner -------------------------------------------------------

@pytest.mark.parametrize("n,shift", [(48, 0.0), (64, 0.3), (1024, 0.03125), (7, 1.0)])
def test_preconditioner_inverts_circulant_laplacian(n, shift):
    kappa, L = 1.3, float(n) * 0.5
    eig = O.precond_eig(n, kappa, L, shift)
    dx = L / n
    sh = shift if shift > 0 else (kappa / dx) * 4 * np.sin(np.pi / n) ** 2
    q = np.random.default_rng(n).standard_normal(n)
    d = O.precondition(q, eig)
    Md = (kappa / dx) * (2 * d - np.roll(d, 1) - np.roll(d, -1)) + sh * d
    np.testing.assert_allclose(Md, q, atol=1e-9 * np.abs(q).max() * max(1.0, (kappa / dx) / sh))


# ---- L-BFGS -----------------------------------------------------------------

def test_lbfgs_golden_fixture():
    # reference tests/test_lbfgs.cpp:309-349
    g = json.load(open(os.path.join(GOLD, "lbfgs_golden.json")))
    nx, ny = int(g["nx"]), int(g["ny"])
    field = np.array(g["field"]["data"]).reshape(nx, ny)
    nz = O.OracleNoise.table(O.prefilter(field), g["Lx"], g["Ly"])
    p = O.params(g["Lx"], g["line_tension"], 0.0, O.LINEAR)
    h, r = O.lbfgs(np.array(g["h0"]["data"]), p, nz, O.lbfgs_params(ftol=1e-8, precondition=True))
    assert r.converged
    e = O.objective(h, p, nz)
    assert abs(e - g["e_min"]) < 1e-5 * abs(g["e_min"]) + 1e-6
    # SciPy's minimiser used a finite-difference gradient; the minimum agrees to ~1e-6
    assert np.max(np.abs(h - np.array(g["h_min"]["data"]))) < 1e-4


def test_lbfgs_relaxes_to_flat_line():
    n = 128
    h0 = 0.5 + 0.3 * np.cos(4 * np.pi * np.arange(n) / n)
    h, r = O.lbfgs(h0, O.params(n, 1.0, 0.0, O.LINEAR), ZERO, O.lbfgs_params(ftol=1e-9))
    assert r.converged and np.ptp(h) < 1e-6


@pytest.mark.parametrize("model", [O.LINEAR, O.ARCLENGTH])
def test_lbfgs_harmonic_minimum(model):
    h, r = O.lbfgs(np.zeros(96), O.params(96, 1.0, 0.0, model), harmonic(2.0, 0.5),
                   O.lbfgs_params(ftol=1e-9))
    assert r.converged and np.max(np.abs(h - 0.5)) < 1e-6


def test_lbfgs_reports_nonconvergence():
    h, r = O.lbfgs(np.zeros(96), O.params(96, 1.0, 0.0, O.LINEAR), harmonic(2.0, 0.5),
                   O.lbfgs_params(ftol=1e-12, max_iter=1))
    assert not r.converged and r.max_force >= 1e-12 and r.iterations == 1


def test_lbfgs_curvature_skip_forced():
    h, r = O.lbfgs(np.zeros(16), O.params(16, 1.0, 0.0, O.LINEAR), harmonic(2.0, 0.5),
                   O.lbfgs_params(ftol=1e-8, max_iter=20000, curvature_eps=1e30))
    assert r.n_curvature_skips > 0 and r.converged and np.max(np.abs(h - 0.5)) < 1e-6


def test_lbfgs_gives_up_without_history():
    h, r = O.lbfgs(np.zeros(32), O.params(32, 1.0, 0.0, O.LINEAR), harmonic(2.0, 0.5),
                   O.lbfgs_params(ftol=1e-9, max_ls=0))
    assert not r.converged and r.iterations == 0 and r.n_ls_restarts == 0


def test_lbfgs_restarts_on_step_failure():
    _, coef = O.filtered_noise(64, 64, 3.2, 1.0, 2.0, 0.15, 0.15, 3)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    p = O.params(3.2, 0.5, 0.0, O.LINEAR)
    _, r = O.lbfgs(np.full(64, 0.5), p, nz, O.lbfgs_params(ftol=1e-6, max_iter=50000, max_ls=3))
    assert r.n_ls_restarts > 0
    _, r2 = O.lbfgs(np.full(64, 0.5), p, nz, O.lbfgs_params(ftol=1e-6, max_iter=50000, max_ls=30))
    assert r2.converged and r2.n_ls_restarts == 0


def test_lbfgs_fast_on_quadratic():
    _, r = O.lbfgs(np.zeros(128), O.params(128, 1.0, 0.0, O.LINEAR), harmonic(1.0, 0.3),
                   O.lbfgs_params(ftol=1e-8))
    assert r.converged and r.iterations < 200


def test_lbfgs_preconditioner_on_stiff_line():
    _, coef = O.filtered_noise(256, 64, 2.56, 1.0, 0.3, 0.2, 0.2, 5)
    nz = O.OracleNoise.table(coef, 2.56, 1.0)
    p = O.params(2.56, 50.0, 0.0, O.LINEAR)
    _, r = O.lbfgs(np.full(256, 0.5), p, nz, O.lbfgs_params(ftol=1e-6, max_iter=5000, precondition=True))
    assert r.converged and r.iterations < 300
    _, r2 = O.lbfgs(np.full(256, 0.5), p, nz, O.lbfgs_params(ftol=1e-6, max_iter=5000))
    assert r.iterations < r2.iterations


# ---- FIRE (src/fire.h; reference tests/test_fire.cpp, tests/test_lbfgs.cpp:246-271) ----

def _fire_params(**kw):
    return O.FireParams(0.1 * math.sqrt(1.0 / 1.0), **kw)


def test_fire_relaxes_to_flat_line():
    n = 128
    h0 = 0.5 + 0.3 * np.cos(4 * np.pi * np.arange(n) / n)
    h, r = O.fire(h0, O.params(n, 1.0, 0.0, O.LINEAR), ZERO, _fire_params(ftol=1e-9))
    assert r.converged and r.max_force < 1e-9 and np.ptp(h) < 1e-6


@pytest.mark.parametrize("model", [O.LINEAR, O.ARCLENGTH])
def test_fire_harmonic_minimum(model):
    h, r = O.fire(np.zeros(96), O.params(96, 1.0, 0.0, model), harmonic(2.0, 0.5),
                  _fire_params(ftol=1e-9))
    assert r.converged and np.max(np.abs(h - 0.5)) < 1e-6


def test_fire_lowers_energy():
    rng = np.random.default_rng(1)
    h0 = rng.random(96)
    p = O.params(96, 1.0, 0.0, O.ARCLENGTH)
    nz = harmonic(1.5, 0.6)
    h, r = O.fire(h0, p, nz, _fire_params(ftol=1e-9))
    assert r.converged and O.objective(h, p, nz) <= O.objective(h0, p, nz) + 1e-12


def test_fire_reports_nonconvergence():
    h, r = O.fire(np.zeros(64), O.params(64, 1.0, 0.0, O.LINEAR), harmonic(2.0, 0.5),
                  _fire_params(ftol=1e-12, max_iter=3))
    assert not r.converged and r.iterations == 3 and r.max_force >= 1e-12


def test_fire_first_step_closed_form():
    """From rest, v = -dt g is parallel to F, so the mixing leaves it
    unchanged and the first step is h1 = h0 - dt^2 g(h0), dt = 0.1 dt_max."""
    rng = np.random.default_rng(4)
    h0 = rng.random(24)
    p = O.params(24, 1.3, 0.05, O.LINEAR)
    nz = harmonic(0.7, 0.2)
    fp = O.FireParams(0.3, ftol=0.0, max_iter=1)
    h, r = O.fire(h0, p, nz, fp)
    dt = 0.03
    expect = h0 - dt * dt * O.gradient(h0, p, nz)
    assert r.iterations == 1 and np.max(np.abs(h - expect)) < 1e-15
    # dt_init, when given, replaces 0.1 dt_max
    h, _ = O.fire(h0, p, nz, O.FireParams(0.3, dt_init=0.2, ftol=0.0, max_iter=1))
    assert np.max(np.abs(h - (h0 - 0.04 * O.gradient(h0, p, nz)))) < 1e-15


def test_fire_agrees_with_lbfgs():
    rng = np.random.default_rng(3)
    h0 = rng.random(96)
    p = O.params(96, 1.0, 0.0, O.LINEAR)
    nz = harmonic(1.5, 0.4)
    ha, _ = O.lbfgs(h0, p, nz, O.lbfgs_params(ftol=1e-9))
    hb, r = O.fire(h0, p, nz, _fire_params(ftol=1e-9))
    assert r.converged and np.max(np.abs(ha - hb)) < 1e-5


# ---- Philox / Langevin --------------------------------------------------------

def test_philox4x32_10_known_answers():
    # Random123 kat_vectors (Salmon et al. 2011), philox4x32 with 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normals_are_standard():
    z = np.array([O.normal(11, s, i) for s in range(40) for i in range(500)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    assert O.normal(11, 3, 5) == O.normal(11, 3, 5)
    assert O.normal(11, 3, 5) != O.normal(12, 3, 5)


def test_langevin_zero_temperature_descent():
    # reference tests/test_dynamics.cpp:44-65
    n = 64
    p = O.params(n, 1.0, 0.0, O.LINEAR)
    nz = harmonic(4.0, 0.0)
    h = np.random.default_rng(1).uniform(-0.5, 0.5, n)
    e_prev = O.objective(h, p, nz)
    for s in range(2000):
        O.langevin_step(h, p, nz, 0.05, 0.0, 0, s)
        e = O.objective(h, p, nz)
        assert e <= e_prev + 1e-12
        e_prev = e
    assert e_prev < 1e-6


def test_langevin_driving_force_shifts_equilibrium():
    # reference tests/test_dynamics.cpp:106-130
    n, k, y0, f = 32, 4.0, 0.2, 0.6
    p = O.params(n, 1.0, f, O.LINEAR)
    h = np.zeros(n)
    for s in range(5000):
        O.langevin_step(h, p, harmonic(k, y0), 0.05, 0.0, 0, s)
    assert np.max(np.abs(h - (y0 + f / k))) < 1e-6


@pytest.mark.slow
def test_langevin_equipartition():
    # reference tests/test_dynamics.cpp:70-99, <E> = (N/2) kT within 10%
    n, drift, kT = 16, 0.02, 0.05
    diff = math.sqrt(2 * drift * kT)
    p = O.params(n, 1.0, 0.0, O.LINEAR)
    nz = harmonic(4.0, 0.0)
    h = np.zeros(n)
    for s in range(20000):
        O.langevin_step(h, p, nz, drift, diff, 2024, s)
    tot = 0.0
    ns = 100000
    for s in range(20000, 20000 + ns):
        O.langevin_step(h, p, nz, drift, diff, 2024, s)
        tot += O.objective(h, p, nz)
    assert abs(tot / ns - 0.5 * n * kT) < 0.1 * 0.5 * n * kT


# ---- Rosso-Krauth -----------------------------------------------------------

def _rk_washboard_params():
    lam = 1.0
    return O.rk_params(dstep=lam / 200, max_advance=lam, root_tol=1e-12, sweep_tol=1e-11,
                       runaway=3 * lam, max_sweeps=100000)


def test_rk_recovers_analytic_critical_force():
    # reference tests/test_rosso_krauth.cpp:49-89
    nz = washboard()
    fc = 0.1 * 2 * math.pi
    rp = _rk_washboard_params()

    def blocks(f):
        return O.rk_block(np.zeros(16), O.params(16, 1.0, f, O.LINEAR), nz, rp).blocked

    lo, hi = 0.0, 2 * fc
    assert blocks(lo) and not blocks(hi)
    for _ in range(10):  # bracket width 2 fc / 1024, far inside the 2% tolerance
        fm = 0.5 * (lo + hi)
        if blocks(fm):
            lo = fm
        else:
            hi = fm
    assert abs(0.5 * (lo + hi) - fc) < 0.02 * fc


@pytest.mark.parametrize("model", [O.LINEAR, O.ARCLENGTH])
def test_rk_site_velocity_equals_minus_gradient(model):
    # reference tests/test_rosso_krauth.cpp:94-141
    _, coef = O.filtered_noise(64, 64, 3.2, 1.0, 0.7, 0.2, 0.2, 13)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    p = O.params(3.2, 1.3, 0.4, model)
    h = np.random.default_rng(21).uniform(0.1, 0.9, 64)
    g = O.gradient(h, p, nz)
    v = [O.rk_site_velocity(i, 64, 3.2 / 64, 1.3, 0.4, model, h[i], h[i - 1], h[(i + 1) % 64], nz)
         for i in range(64)]
    assert np.max(np.abs(np.array(v) + g)) < 1e-12


@pytest.mark.parametrize("model,amp,lt,f,seed", [(O.LINEAR, 0.5, 1.0, 0.0, 7),
                                                 (O.ARCLENGTH, 0.6, 1.2, 0.1, 29)])
def test_rk_blocked_config_forward_and_arrested(model, amp, lt, f, seed):
    # reference tests/test_rosso_krauth.cpp:145-186, 255-291
    _, coef = O.filtered_noise(64, 64, 3.2, 1.0, amp, 0.2, 0.2, seed)
    nz = O.OracleNoise.table(coef, 3.2, 1.0)
    p = O.params(3.2, lt, f, model)
    h = np.full(64, 0.5)
    rp = O.rk_params(dstep=0.5 / 64, max_advance=1.0, root_tol=1e-11, sweep_tol=1e-9, runaway=3.0)
    r = O.rk_block(h, p, nz, rp)
    assert r.blocked and np.all(h >= 0.5 - 1e-12)
    assert np.max(-O.gradient(h, p, nz)) < 1e-6


def test_rk_velocity_evaluation_count():
    # One sweep of a flat line in zero potential, linear, dx = lt = 1, f = 0.25:
    # a site with fixed neighbours at height c stops at c + f/2. Colour 0
    # (neighbours at 0): v0 > 0, marches to 0.1 (v > 0) and 0.2 (v < 0), then
    # bisects [0.1, 0.2] down to 0.1/2^7 < 1e-3: 1 + 2 + 7 = 10 evaluations.
    # Colour 1 (neighbours at ~0.125) stops near 0.25: marches 0.1, 0.2, 0.3,
    # bisects [0.2, 0.3]: 1 + 3 + 7 = 11. Total 8*10 + 8*11.
    rp = O.rk_params(dstep=0.1, max_advance=10.0, root_tol=1e-3, sweep_tol=1e-12,
                     runaway=100.0, max_sweeps=1)
    h = np.zeros(16)
    r = O.rk_block(h, O.params(16, 1.0, 0.25, O.LINEAR), O.OracleNoise(O.ZERO), rp)
    assert r.sweeps == 1 and not r.blocked and not r.runaway
    assert r.n_velocity_evals == 8 * 10 + 8 * 11
    assert np.all(np.abs(h[0::2] - 0.125) < 1e-3) and np.all(np.abs(h[1::2] - 0.25) < 2e-3)
    # a line at rest: one evaluation per site, blocked after one sweep
    r = O.rk_block(np.zeros(16), O.params(16, 1.0, 0.0, O.LINEAR), O.OracleNoise(O.ZERO), rp)
    assert r.blocked and r.n_velocity_evals == 16


def test_rk_runaway_flag_and_odd_n():
    r = O.rk_block(np.zeros(16), O.params(16, 1.0, 2 * 0.2 * math.pi, O.LINEAR), washboard(),
                   _rk_washboard_params())
    assert r.runaway and not r.blocked and r.mean_h > 3.0
    with pytest.raises(ValueError):
        O.rk_block(np.zeros(17), O.params(17, 1.0, 0.0, O.LINEAR), washboard(),
                   O.rk_params(dstep=0.01, max_advance=1.0, runaway=3.0))
