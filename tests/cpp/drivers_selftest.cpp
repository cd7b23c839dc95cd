// GPU checks of the C++ facade and the run drivers (include/qew/), restating
// the reference's driver tests (tests/test_executables.cpp:50-186) and the
// facade's reference semantics (odd-length rk_block throws
// std::invalid_argument, tests/test_rosso_krauth.cpp:296-313). Prints one
// line per failed check; exit status = number of failures (capped at 1).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "qew/qew_io.hpp"
#include "qew/qew_runs.hpp"

using namespace qew;

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

static void fmt_specs() {
    CHECK(fmt("%.1g", 0.0) == "0");
    CHECK(fmt("%.3g", 0.1) == "0.1");
    CHECK(fmt("%.3g", 100.0) == "100");
    CHECK(fmt("%.6g", 1.0) == "1");
    CHECK(fmt("%.4g", 0.8) == "0.8");
}

static void static_sweep_shape_and_mapping() {
    StaticConfig cfg;
    cfg.model = Model::Linear;
    cfg.amplitude = 1.3;
    cfg.xi = 0.1;
    cfg.Lx = 1.0;
    cfg.Ly = 1.0;
    cfg.nx = 16;
    cfg.ny = 8;
    cfg.pinning_lengths_over_xi = {0.1, 1.0, 4.0};
    cfg.ftol = 1e-8;
    cfg.max_iter = 2000;
    const auto profiles = static_sweep(cfg, ZeroNoise{});
    CHECK(profiles.size() == 3);
    for (size_t i = 0; i < profiles.size(); ++i) {
        const auto &prof = profiles[i];
        CHECK(prof.pinning_length_over_xi == cfg.pinning_lengths_over_xi[i]);
        CHECK(std::abs(prof.line_tension -
                       std::pow(cfg.pinning_lengths_over_xi[i], 1.5) * cfg.amplitude) < 1e-12);
        CHECK(prof.h.size() == 16u);
        CHECK(prof.result.converged);
        for (double v : prof.h) CHECK(std::abs(v - 0.5) < 1e-9);
    }
}

// The batched sweep (all pinning lengths in one persistent launch) reaches the
// minima of the reference's one-length-at-a-time sweep; a line length the
// batched preconditioner does not take falls back to the sequential sweep.
static void static_sweep_batched_equals_sequential() {
    StaticConfig cfg;
    cfg.model = Model::Linear;
    cfg.xi = 0.1;
    cfg.Lx = 2.56;
    cfg.Ly = 1.0;
    cfg.nx = 256;
    cfg.ny = 64;
    cfg.pinning_lengths_over_xi = {1.0, 4.0, 10.0};
    cfg.ftol = 1e-7;  // above the energy roundoff floor of these lines
    cfg.max_iter = 5000;
    FilteredNoise noise(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, 3);
    cfg.batched = true;
    const auto b = static_sweep(cfg, noise);
    cfg.batched = false;
    const auto s = static_sweep(cfg, noise);
    CHECK(b.size() == 3 && s.size() == 3);
    for (size_t k = 0; k < b.size() && k < s.size(); ++k) {
        CHECK(b[k].result.converged && s[k].result.converged);
        CHECK(b[k].line_tension == s[k].line_tension);
        CHECK(b[k].h.size() == 256u && s[k].h.size() == 256u);
        double dh = 0;
        for (size_t i = 0; i < b[k].h.size() && i < s[k].h.size(); ++i)
            dh = std::max(dh, std::abs(b[k].h[i] - s[k].h[i]));
        CHECK(dh < 1e-5);
    }
    // the lengths the batched kernel takes: even, <= 16384, half 7-smooth
    CHECK(batch_lbfgs_supports(24, 8) && batch_lbfgs_supports(210, 8));
    CHECK(batch_lbfgs_supports(16384, 32) && batch_lbfgs_supports(11250, 10));
    CHECK(!batch_lbfgs_supports(22, 8) && !batch_lbfgs_supports(25, 8));
    CHECK(!batch_lbfgs_supports(2, 8) && !batch_lbfgs_supports(16386, 8));
    CHECK(!batch_lbfgs_supports(256, 33) && !batch_lbfgs_supports(256, 0));
    StaticConfig odd;
    odd.model = Model::Linear;
    odd.Lx = 1.0;
    odd.nx = 22;  // half is 11, a prime the LDS FFT has no stage for: the sequential path
    odd.pinning_lengths_over_xi = {0.5, 2.0};
    odd.ftol = 1e-8;
    const auto o = static_sweep(odd, ZeroNoise{});
    CHECK(o.size() == 2);
    for (const auto &prof : o) {
        CHECK(prof.result.converged && prof.h.size() == 22u);
        for (double v : prof.h) CHECK(std::abs(v - 0.5) < 1e-9);
    }
}

static void dynamic_anneal_timestep_and_snapshots() {
    DynamicConfig cfg;
    cfg.model = Model::Linear;
    cfg.amplitude = 1.0;
    cfg.xi = 0.1;
    cfg.pinning_length_over_xi = 0.1;
    cfg.Lx = 1.0;
    cfg.Ly = 1.0;
    cfg.nx = 32;
    cfg.ny = 8;
    cfg.tau = 1.0;
    cfg.tl_start = 0.01;
    cfg.tl_end = 0.0001;
    cfg.total_time = 200.0;
    const DynamicResult res = dynamic_anneal(cfg, ZeroNoise{});
    const double fac = cfg.xi / std::max(cfg.tl_start, cfg.tl_end);
    CHECK(std::abs(res.dt - std::min(0.1, cfg.tau / 10 * std::min(fac, fac * fac))) < 1e-12);
    CHECK(std::abs(res.line_tension - std::pow(0.1, 1.5)) < 1e-12);
    CHECK(res.snapshots.size() >= 2);
    CHECK(std::abs(res.snapshots.front().t - 10 * res.dt) <= res.dt);
    for (size_t i = 1; i < res.snapshots.size(); ++i) {
        CHECK(res.snapshots[i].t > res.snapshots[i - 1].t);
        CHECK(std::abs(res.snapshots[i].t / res.snapshots[i - 1].t - 10.0) < 0.5);
    }
    for (const auto &snap : res.snapshots) {
        CHECK(snap.h.size() == 32u);
        for (double v : snap.h) CHECK(std::isfinite(v));
    }
    CHECK(res.snapshots.back().t <= res.total_time);
}

// dynamic_anneal sends the steps between snapshots to the device together;
// the snapshots must equal the reference loop run one step per call
static void dynamic_anneal_matches_step_loop() {
    DynamicConfig cfg;
    cfg.model = Model::Arclength;
    cfg.amplitude = 1.0;
    cfg.xi = 0.1;
    cfg.pinning_length_over_xi = 1.0;
    cfg.Lx = 3.0;
    cfg.Ly = 1.0;
    cfg.nx = 300;
    cfg.ny = 64;
    cfg.seed = 5;
    cfg.tl_start = 0.01;
    cfg.tl_end = 0.001;
    cfg.total_time = 50.0;
    const FilteredNoise noise(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, 7);
    const DynamicResult res = dynamic_anneal(cfg, noise);

    const double dt = dynamic_timestep(cfg);
    Params p;
    p.physical_size = cfg.Lx;
    p.line_tension = larkin_line_tension(cfg.pinning_length_over_xi, cfg.amplitude);
    p.driving_force = cfg.driving_force;
    p.model = cfg.model;
    View1D h("h", cfg.nx);
    deep_copy(h, 0.5 * cfg.Ly);
    LangevinDynamics dyn(cfg.nx, cfg.seed + 1);
    double t = 0, next_save = 10 * dt, tl = cfg.tl_start;
    size_t k = 0;
    while (t < cfg.total_time) {
        dyn.step(h, p, noise, tl * dt / (cfg.tau * cfg.amplitude), tl * std::sqrt(2 * dt / cfg.tau));
        tl = cfg.tl_start + (cfg.tl_end - cfg.tl_start) * (t / cfg.total_time);
        t += dt;
        if (t > next_save - dt / 2) {
            CHECK(k < res.snapshots.size());
            if (k < res.snapshots.size()) {
                CHECK(res.snapshots[k].t == t);
                CHECK(res.snapshots[k].h == to_host_vec(h));
            }
            ++k;
            next_save *= 10;
        }
    }
    CHECK(k == res.snapshots.size() && k >= 2);
}

static void depinning_ramp_brackets_critical_force() {
    const SinusoidNoise noise;
    const double fc_exact = noise.A * noise.k;
    const double lambda = 2 * 3.14159265358979323846 / noise.k;
    DepinningConfig cfg;
    cfg.model = Model::Linear;
    cfg.amplitude = 1.0;
    cfg.pinning_length_over_xi = 1.0;
    cfg.Lx = 8.0;
    cfg.Ly = 1.0;
    cfg.nx = 8;
    cfg.ny = 8;
    cfg.f_step = 0.2;
    cfg.f_max = 2.0;
    cfg.rp.dstep = lambda / 100;
    cfg.rp.max_advance = lambda;
    cfg.rp.runaway = 3 * lambda;
    cfg.rp.root_tol = 1e-11;
    cfg.rp.sweep_tol = 1e-8;
    cfg.rp.max_sweeps = 20000;
    const DepinningResult res = depinning_ramp(cfg, noise);
    CHECK(std::abs(res.line_tension - 1.0) < 1e-12);
    CHECK(!res.steps.empty());
    CHECK(res.steps.front().result.blocked);
    CHECK(res.steps.front().f == 0.0);
    const DepinningStep &last = res.steps.back();
    CHECK(!last.result.blocked && last.result.runaway);
    CHECK(last.h.empty());
    CHECK(std::abs(res.f_c_upper_bound - last.f) < 1e-9);
    CHECK(res.f_c_upper_bound - cfg.f_step <= fc_exact);
    CHECK(fc_exact <= res.f_c_upper_bound + 1e-9);
    for (const auto &st : res.steps)
        if (st.result.blocked) CHECK(st.h.size() == 8u);

    // bisection from the ramp's bracket (warm-started probes)
    cfg.fc_rel_tol = 1e-6;
    const DepinningResult rb = depinning_ramp(cfg, noise);
    CHECK(rb.f_c_lower <= rb.f_c_upper);
    CHECK(rb.f_c_upper - rb.f_c_lower <= 1e-6 * rb.f_c_upper * 1.0000001);
    CHECK(std::abs(rb.f_c - fc_exact) < 0.02 * fc_exact);
    CHECK(rb.h_critical.size() == 8u);

    // four forces per round: the same critical force, fewer rounds
    cfg.fc_probes = 4;
    const DepinningResult r4 = depinning_ramp(cfg, noise);
    CHECK(r4.f_c_upper - r4.f_c_lower <= 1e-6 * r4.f_c_upper * 1.0000001);
    CHECK(r4.f_c_lower <= rb.f_c_upper && rb.f_c_lower <= r4.f_c_upper);  // brackets overlap
    CHECK(r4.bisection_probes % 4 == 0 && r4.bisection_probes / 4 < rb.bisection_probes);
    CHECK(r4.h_critical.size() == 8u);
}

static void facade_semantics() {
    // odd n: std::invalid_argument, as the reference throws
    View1D h("h", 17);
    h.fill(0.0);
    Params p;
    p.physical_size = 17;
    RkParams rp;
    rp.dstep = 0.01;
    rp.max_advance = 1.0;
    rp.runaway = 3.0;
    bool threw = false;
    try {
        rk_block(h, p, SinusoidNoise{}, rp);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw);
    // objective of a flat line in a constant potential (tests/test_qew_model.cpp:110-132)
    View1D f("f", 50);
    f.fill(0.7);
    Params q;
    q.physical_size = 3.0;
    q.line_tension = 2.3;
    q.driving_force = 0.4;
    CHECK(std::abs(objective(f, q, ConstNoise{1.25}) - 3.0 * (1.25 - 0.4 * 0.7)) < 1e-12);
    // L-BFGS reaches the harmonic minimum (tests/test_lbfgs.cpp:87-102)
    View1D w("w", 96);
    w.fill(0.0);
    Params lp;
    lp.physical_size = 96;
    LbfgsParams opt;
    opt.ftol = 1e-9;
    const LbfgsResult r = lbfgs_minimize(w, lp, HarmonicWell{2.0, 0.5}, opt);
    CHECK(r.converged);
    for (double v : w.to_host()) CHECK(std::abs(v - 0.5) < 1e-6);
    // FIRE reaches the same minimum (tests/test_fire.cpp:82-108)
    View1D fw("fw", 96);
    fw.fill(0.0);
    FireParams fopt;
    fopt.dt_max = 0.1;
    fopt.ftol = 1e-9;
    const FireResult fr = fire_minimize(fw, lp, HarmonicWell{2.0, 0.5}, fopt);
    CHECK(fr.converged && fr.max_force < 1e-9);
    for (double v : fw.to_host()) CHECK(std::abs(v - 0.5) < 1e-6);
    bool fire_threw = false;
    try {
        fire_minimize(fw, lp, HarmonicWell{2.0, 0.5}, FireParams{});  // dt_max missing
    } catch (const std::invalid_argument &) {
        fire_threw = true;
    }
    CHECK(fire_threw);
    // LangevinDynamics at zero diffusion is seed independent (tests/test_dynamics.cpp:137-161)
    View1D a("a", 48), b("b", 48);
    a.fill(0.3);
    b.fill(0.3);
    LangevinDynamics d1(48, 111), d2(48, 999);
    Params dp;
    dp.physical_size = 48;
    dp.driving_force = 0.3;
    for (int s = 0; s < 50; ++s) {
        d1.step(a, dp, HarmonicWell{4.0, 0.1}, 0.05, 0.0);
        d2.step(b, dp, HarmonicWell{4.0, 0.1}, 0.05, 0.0);
    }
    CHECK(a.to_host() == b.to_host());
}

int main() {
    fmt_specs();
    static_sweep_shape_and_mapping();
    static_sweep_batched_equals_sequential();
    dynamic_anneal_timestep_and_snapshots();
    dynamic_anneal_matches_step_loop();
    depinning_ramp_brackets_critical_force();
    facade_semantics();
    finalize();
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
