// Run drivers behind the command-line tools: the reference's static sweep,
// annealed dynamics and force ramp (src/qew_runs.h:45-245), plus the f_c
// bisection the depinning workload needs (absent from the reference drivers;
// its only bisection is tests/test_rosso_krauth.cpp:76-88).
//
// Kept identical to the reference because the outputs must match:
//   * line tension from lambda_p / xi: Gamma = A (lambda_p/xi)^{3/2}
//     (src/qew_runs.h:77,143-144,210-211);
//   * flat start h = Ly / 2;
//   * static: preconditioned L-BFGS with shift 0.5 dx A / xi^2 (:87-91);
//   * dynamic: dt = min(0.1, tau/10 min(fac, fac^2)), fac = xi / max(tl_start,
//     tl_end) (:139-142); drift tl dt/(tau A), diffusion tl sqrt(2 dt/tau)
//     (:165-166); tl updated after each step from the pre-increment time
//     (:167-170); snapshots at 10 dt, 100 dt, ... (:171-174); Langevin seed is
//     seed + 1 (:154);
//   * ramp: f = k f_step with an integer k, warm start from the previous
//     configuration, stop at the first non-blocked force (:224-241).
// The static sweep's pinning lengths are independent minimisations of one
// disorder, so they run together in the persistent batched L-BFGS kernel
// (the reference minimises them one after the other; StaticConfig::batched
// = false does the same here).
#ifndef QEW_RUNS_HPP
#define QEW_RUNS_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#include "qew.hpp"

namespace qew {

struct StaticConfig {
    Model model = Model::Arclength;
    real_t amplitude = 1.0;
    real_t xi = 0.1;
    real_t Lx = 32.0;
    real_t Ly = 1.0;
    int nx = 4096;
    int ny = 512;
    real_t driving_force = 0.0;
    std::vector<double> pinning_lengths_over_xi = {0.01, 0.1, 1.0, 10.0, 100.0};
    real_t ftol = 1e-6;
    int max_iter = 5000;
    int m = 8;  // history length (the reference fixes the LbfgsParams default)
    // true: the pinning lengths are minimised together by the persistent
    // batched L-BFGS kernel (one workgroup per length) when it supports the
    // line (even nx in [4, 16384] whose half has no prime factor above 7,
    // m <= 32); false, or an
    // unsupported line: one host-driven minimisation per length, as the
    // reference does. No reference counterpart.
    bool batched = true;
};

struct StaticProfile {
    double pinning_length_over_xi = 0;
    real_t line_tension = 0;
    std::vector<double> h;
    LbfgsResult result;
};

inline real_t larkin_line_tension(double lp_over_xi, real_t amplitude) {
    return std::pow(lp_over_xi, static_cast<real_t>(1.5)) * amplitude;
}

// True when the persistent batched L-BFGS kernel accepts a preconditioned
// line of n sites with history m: its FFT transforms the n/2 complex values of
// the packed real line in one workgroup's LDS with radix 2, 3, 4, 5 and 7
// stages, so n is even, at most 16384, and n/2 has no other prime factor.
inline bool batch_lbfgs_supports(int n, int m) {
    if (n < 4 || n > 16384 || n % 2 != 0 || m < 1 || m > 32) return false;
    int r = n / 2;
    for (int f : {2, 3, 5, 7})
        while (r % f == 0) r /= f;
    return r == 1;
}

template <class Noise>
std::vector<StaticProfile> static_sweep(const StaticConfig &c, const Noise &noise) {
    LbfgsParams opt;
    opt.m = c.m;
    opt.ftol = c.ftol;
    opt.max_iter = c.max_iter;
    opt.precondition = true;
    opt.precond_shift = static_cast<real_t>(0.5) * (c.Lx / c.nx) * c.amplitude / (c.xi * c.xi);
    const size_t np = c.pinning_lengths_over_xi.size();
    std::vector<Params> ps(np);
    std::vector<StaticProfile> out(np);
    for (size_t k = 0; k < np; ++k) {
        const double lp = c.pinning_lengths_over_xi[k];
        ps[k].physical_size = c.Lx;
        ps[k].line_tension = larkin_line_tension(lp, c.amplitude);
        ps[k].driving_force = c.driving_force;
        ps[k].model = c.model;
        out[k].pinning_length_over_xi = lp;
        out[k].line_tension = ps[k].line_tension;
    }
    if (np == 0) return out;
    if (c.batched && batch_lbfgs_supports(c.nx, c.m)) {
        // every length's line back to back in one buffer, one launch
        const DeviceNoise dn = as_device(noise);
        std::vector<qew_noise> nh(np, dn.handle());
        std::vector<int> len(np, c.nx);
        std::vector<qew_params> cp(np);
        for (size_t k = 0; k < np; ++k) cp[k] = ps[k].c();
        qew_batch b = nullptr;
        check(qew_batch_create(context(), (int)np, len.data(), nh.data(), cp.data(), &b));
        std::unique_ptr<qew_batch_s, qew_status (*)(qew_batch)> owner(b, qew_batch_destroy);
        View1D h("h", np * (size_t)c.nx);
        deep_copy(h, static_cast<real_t>(0.5) * c.Ly);
        const qew_lbfgs_params l = opt.c();
        std::vector<qew_lbfgs_result> r(np);
        check(qew_batch_lbfgs(b, QEW_F64, &l, h.data(), r.data()));
        const std::vector<double> all = to_host_vec(h);
        for (size_t k = 0; k < np; ++k) {
            out[k].result = from_c(r[k]);
            out[k].h.assign(all.begin() + k * (size_t)c.nx, all.begin() + (k + 1) * (size_t)c.nx);
        }
        return out;
    }
    for (size_t k = 0; k < np; ++k) {
        View1D h("h", c.nx);
        deep_copy(h, static_cast<real_t>(0.5) * c.Ly);
        out[k].result = lbfgs_minimize(h, ps[k], noise, opt);
        out[k].h = to_host_vec(h);
    }
    return out;
}

struct DynamicConfig {
    Model model = Model::Arclength;
    real_t amplitude = 1.0;
    real_t xi = 0.1;
    real_t pinning_length_over_xi = 0.1;
    real_t driving_force = 0.0;
    real_t Lx = 2.0;
    real_t Ly = 1.0;
    int nx = 2048;
    int ny = 1024;
    std::uint64_t seed = 1;
    real_t tau = 1.0;
    real_t tl_start = 0.01;
    real_t tl_end = 0.0001;
    real_t total_time = 1000.0;
    real_t dt = 0;  // > 0 overrides the timestep heuristic (no reference counterpart)
};

struct DynamicSnapshot {
    real_t t = 0;
    std::vector<double> h;
};

struct DynamicResult {
    real_t dt = 0;
    real_t total_time = 0;
    real_t line_tension = 0;
    std::vector<DynamicSnapshot> snapshots;
};

inline real_t dynamic_timestep(const DynamicConfig &c) {
    if (c.dt > 0) return c.dt;
    const real_t fac = c.xi / std::max(c.tl_start, c.tl_end);
    return std::min(static_cast<real_t>(0.1), c.tau / 10 * std::min(fac, fac * fac));
}

template <class Noise>
DynamicResult dynamic_anneal(const DynamicConfig &c, const Noise &noise) {
    const real_t tau = c.tau;
    const real_t dt = dynamic_timestep(c);
    const real_t line_tension = larkin_line_tension(c.pinning_length_over_xi, c.amplitude);
    Params p;
    p.physical_size = c.Lx;
    p.line_tension = line_tension;
    p.driving_force = c.driving_force;
    p.model = c.model;
    View1D h("h", c.nx);
    deep_copy(h, static_cast<real_t>(0.5) * c.Ly);
    LangevinDynamics dyn(c.nx, c.seed + 1);
    const DeviceNoise dn = as_device(noise);
    DynamicResult res;
    res.dt = dt;
    res.total_time = c.total_time;
    res.line_tension = line_tension;
    real_t t = 0;
    real_t next_save = 10 * dt;
    real_t tl = c.tl_start;
    // the schedule is run on the host exactly as the reference loop runs it;
    // the steps between snapshots go to the device together
    std::vector<real_t> drift, diff;
    auto flush = [&] {
        dyn.steps(h, p, dn, drift.data(), diff.data(), (int)drift.size());
        drift.clear();
        diff.clear();
    };
    while (t < c.total_time) {
        drift.push_back(tl * dt / (tau * c.amplitude));
        diff.push_back(tl * std::sqrt(2 * dt / tau));
        tl = c.tl_start + (c.tl_end - c.tl_start) * (t / c.total_time);
        t += dt;
        if (t > next_save - dt / 2) {
            flush();
            res.snapshots.push_back(DynamicSnapshot{t, to_host_vec(h)});
            next_save *= 10;
        } else if (drift.size() >= (1u << 16)) {
            flush();
        }
    }
    flush();
    return res;
}

struct DepinningConfig {
    Model model = Model::Arclength;
    real_t amplitude = 1.0;
    real_t xi = 0.1;
    real_t pinning_length_over_xi = 5.0;
    real_t Lx = 8.0;
    real_t Ly = 1.0;
    int nx = 1024;
    int ny = 256;
    real_t f_step = 1.0;
    real_t f_max = 40.0;
    RkParams rp;
    // > 0: after the ramp brackets f_c, bisect [last blocked, first unblocked]
    // to this relative width, warm-starting every probe from the highest
    // blocked configuration found so far (no-passing keeps it below the
    // critical line). No reference counterpart.
    real_t fc_rel_tol = 0;
    // forces probed at once per round of that search: 1 bisects; K > 1 probes
    // lo + j (hi - lo) / (K + 1), j = 1..K, as one batch of K copies of the
    // line and keeps the bracket between the highest force below which every
    // probe blocked and the next probe (K + 1 times narrower per round).
    int fc_probes = 1;
};

struct DepinningStep {
    real_t f = 0;
    RkResult result;
    std::vector<double> h;
};

struct DepinningResult {
    real_t line_tension = 0;
    real_t f_c_upper_bound = -1;
    std::vector<DepinningStep> steps;
    // bisection (fc_rel_tol > 0 and a bracket was found)
    real_t f_c = -1;
    real_t f_c_lower = -1;
    real_t f_c_upper = -1;
    int bisection_probes = 0;
    std::vector<double> h_critical;  // blocked configuration at f_c_lower
};

template <class Noise>
DepinningResult depinning_ramp(const DepinningConfig &c, const Noise &noise) {
    const real_t line_tension = larkin_line_tension(c.pinning_length_over_xi, c.amplitude);
    Params p;
    p.physical_size = c.Lx;
    p.line_tension = line_tension;
    p.model = c.model;
    View1D h("h", c.nx);
    deep_copy(h, static_cast<real_t>(0.5) * c.Ly);
    const DeviceNoise dn = as_device(noise);
    DepinningResult res;
    res.line_tension = line_tension;
    std::vector<double> last_blocked;
    real_t f_lo = -1;
    for (int k = 0; static_cast<real_t>(k) * c.f_step <= c.f_max; ++k) {
        const real_t f = static_cast<real_t>(k) * c.f_step;
        p.driving_force = f;
        const RkResult r = rk_block(h, p, dn, c.rp);
        DepinningStep st;
        st.f = f;
        st.result = r;
        if (r.blocked) {
            st.h = to_host_vec(h);
            last_blocked = st.h;
            f_lo = f;
        } else {
            res.f_c_upper_bound = f;
        }
        res.steps.push_back(std::move(st));
        if (!r.blocked) break;
    }
    if (c.fc_rel_tol > 0 && res.f_c_upper_bound >= 0 && f_lo >= 0 && c.fc_probes > 1) {
        real_t lo = f_lo, hi = res.f_c_upper_bound;
        std::vector<double> h_lo = last_blocked;
        const int K = c.fc_probes;
        const size_t n = (size_t)c.nx;
        std::vector<int> lens(K, c.nx);
        std::vector<qew_noise> nzs(K, dn.handle());
        std::vector<qew_params> ps(K, p.c());
        qew_batch b = nullptr;
        check(qew_batch_create(context(), K, lens.data(), nzs.data(), ps.data(), &b));
        std::unique_ptr<qew_batch_s, qew_status (*)(qew_batch)> owner(b, qew_batch_destroy);
        View1D hb("h_probes", K * n);
        std::vector<double> hh(K * n), forces(K);
        std::vector<qew_rk_result> rr(K);
        const qew_rk_params rpc = c.rp.c();
        while (hi - lo > c.fc_rel_tol * std::max(std::abs(hi), std::abs(lo))) {
            bool distinct = true;
            for (int j = 0; j < K; ++j) {
                forces[j] = lo + (j + 1) * (hi - lo) / (K + 1);
                distinct = distinct && forces[j] > (j ? forces[j - 1] : lo) && forces[j] < hi;
                std::copy(h_lo.begin(), h_lo.end(), hh.begin() + j * n);
            }
            if (!distinct) break;  // bracket at FP64 resolution
            hb.copy_from_host(hh.data());
            check(qew_batch_set_forces(b, forces.data()));
            check(qew_batch_rk_block(b, &rpc, hb.data(), rr.data()));
            res.bisection_probes += K;
            int jb = 0;  // probes 1..jb all blocked
            while (jb < K && rr[jb].blocked) ++jb;
            if (jb > 0) {
                lo = forces[jb - 1];
                hb.copy_to_host(hh.data());
                h_lo.assign(hh.begin() + (jb - 1) * n, hh.begin() + jb * n);
            }
            if (jb < K) hi = forces[jb];
        }
        res.f_c_lower = lo;
        res.f_c_upper = hi;
        res.f_c = 0.5 * (lo + hi);
        res.h_critical = h_lo;
    } else if (c.fc_rel_tol > 0 && res.f_c_upper_bound >= 0 && f_lo >= 0) {
        real_t lo = f_lo, hi = res.f_c_upper_bound;
        std::vector<double> h_lo = last_blocked;
        while (hi - lo > c.fc_rel_tol * std::max(std::abs(hi), std::abs(lo))) {
            const real_t mid = 0.5 * (lo + hi);
            if (!(mid > lo && mid < hi)) break;  // bracket at FP64 resolution
            deep_copy(h, h_lo);
            p.driving_force = mid;
            const RkResult r = rk_block(h, p, dn, c.rp);
            ++res.bisection_probes;
            if (r.blocked) {
                lo = mid;
                h_lo = to_host_vec(h);
            } else {
                hi = mid;
            }
        }
        res.f_c_lower = lo;
        res.f_c_upper = hi;
        res.f_c = 0.5 * (lo + hi);
        res.h_critical = h_lo;
    }
    return res;
}

}  // namespace qew

#endif  // QEW_RUNS_HPP
