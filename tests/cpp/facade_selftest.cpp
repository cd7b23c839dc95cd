// The facade's per-point API on the GPU, restating the reference's tests that
// call DeviceNoise::sample and rk_site_velocity from their own kernels
// (tests/test_filtered_noise.cpp:118-281, tests/test_rosso_krauth.cpp:94-141),
// plus the GradientWorkspace's per-edge arrays and rk_forward_zero. Prints one
// line per failed check and exits non-zero if any failed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <utility>
#include <vector>

#include "qew/qew.hpp"

using namespace qew;

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

// tests/test_filtered_noise.cpp:118-144: the interpolant reproduces the
// field at the grid nodes
static void reproduces_values_at_nodes() {
    const int nx = 64, ny = 48;
    FilteredNoise noise(nx, ny, 3.2, 1.0, 1.5, 0.15, 0.15, 3);
    const std::vector<real_t> f = noise.field();
    std::vector<real_t> xs, ys;
    std::vector<std::pair<int, int>> nodes;
    for (int i = 0; i < nx; i += 5)
        for (int j = 0; j < ny; j += 5) {
            xs.push_back(i * noise.dx());
            ys.push_back(j * noise.dy());
            nodes.emplace_back(i, j);
        }
    const std::vector<NoiseSample> s = noise.device_noise().sample(xs, ys);
    double err = 0;
    for (size_t k = 0; k < nodes.size(); ++k)
        err = std::max(err, std::abs(s[k].v - f[(size_t)nodes[k].first * ny + nodes[k].second]));
    CHECK(err < 1e-9);
}

// tests/test_filtered_noise.cpp:147-181: analytic derivatives against central
// differences of the value
static void derivatives_match_finite_difference() {
    FilteredNoise noise(64, 48, 3.2, 1.0, 1.5, 0.15, 0.15, 5);
    const DeviceNoise dn = noise.device_noise();
    const real_t ex = 1e-5 * noise.dx(), ey = 1e-5 * noise.dy();
    std::mt19937 rng(99);
    std::uniform_real_distribution<real_t> ux(0, 3.2), uy(0, 1.0);
    std::vector<real_t> xs, ys;
    for (int k = 0; k < 100; ++k) {
        xs.push_back(ux(rng));
        ys.push_back(uy(rng));
    }
    std::vector<real_t> xp = xs, xm = xs, yp = ys, ym = ys;
    for (size_t k = 0; k < xs.size(); ++k) {
        xp[k] += ex;
        xm[k] -= ex;
        yp[k] += ey;
        ym[k] -= ey;
    }
    const auto s0 = dn.sample(xs, ys), sxp = dn.sample(xp, ys), sxm = dn.sample(xm, ys);
    const auto syp = dn.sample(xs, yp), sym = dn.sample(xs, ym);
    double mdx = 0, mdy = 0;
    for (size_t k = 0; k < xs.size(); ++k) {
        mdx = std::max(mdx, std::abs((sxp[k].v - sxm[k].v) / (2 * ex) - s0[k].dv_dx));
        mdy = std::max(mdy, std::abs((syp[k].v - sym[k].v) / (2 * ey) - s0[k].dv_dy));
    }
    CHECK(mdx < 1e-5);
    CHECK(mdy < 1e-5);
    // the value-only and dy-only samplers give the same numbers as sample()
    for (int k = 0; k < 5; ++k) {
        CHECK(dn.sample_value(xs[k], ys[k]) == s0[k].v);
        CHECK(dn.sample_dy(xs[k], ys[k]) == s0[k].dv_dy);
        const NoiseSample one = dn.sample(xs[k], ys[k]);
        CHECK(one.v == s0[k].v && one.dv_dx == s0[k].dv_dx && one.dv_dy == s0[k].dv_dy);
    }
}

// tests/test_filtered_noise.cpp:238-281: periodic in both directions, many
// periods out and negative coordinates included
static void periodic_wrap_across_many_cells() {
    const real_t Lx = 2.5, Ly = 1.3;
    FilteredNoise noise(48, 40, Lx, Ly, 1.0, 0.25, 0.25, 17);
    const DeviceNoise dn = noise.device_noise();
    std::mt19937 rng(123);
    std::uniform_real_distribution<real_t> ux(0, Lx), uy(0, Ly);
    std::vector<real_t> xs, ys;
    for (int k = 0; k < 60; ++k) {
        xs.push_back(ux(rng));
        ys.push_back(uy(rng));
    }
    const auto s0 = dn.sample(xs, ys);
    const std::pair<int, int> shifts[] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1},
                                          {3, -2}, {-5, 7}, {-13, -11}};
    for (const auto &sh : shifts) {
        std::vector<real_t> xss = xs, yss = ys;
        for (size_t k = 0; k < xs.size(); ++k) {
            xss[k] += sh.first * Lx;
            yss[k] += sh.second * Ly;
        }
        const auto s = dn.sample(xss, yss);
        double mv = 0, mdx = 0, mdy = 0;
        for (size_t k = 0; k < xs.size(); ++k) {
            mv = std::max(mv, std::abs(s[k].v - s0[k].v));
            mdx = std::max(mdx, std::abs(s[k].dv_dx - s0[k].dv_dx));
            mdy = std::max(mdy, std::abs(s[k].dv_dy - s0[k].dv_dy));
        }
        CHECK(mv < 1e-9 && mdx < 1e-9 && mdy < 1e-9);
    }
}

// tests/test_rosso_krauth.cpp:94-141: the site velocity with neighbours fixed
// is -grad_i for arbitrary configurations
static void site_velocity_matches_gradient() {
    const int nx = 64;
    const real_t Lx = 3.2;
    FilteredNoise noise(nx, 64, Lx, 1.0, 0.7, 0.2, 0.2, 13);
    const DeviceNoise dn = noise.device_noise();
    for (const Model model : {Model::Linear, Model::Arclength}) {
        Params p;
        p.physical_size = Lx;
        p.line_tension = 1.3;
        p.driving_force = 0.4;
        p.model = model;
        std::mt19937 rng(21);
        std::uniform_real_distribution<real_t> dist(0.1, 0.9);
        std::vector<real_t> hh(nx), hl(nx), hr(nx);
        for (int i = 0; i < nx; ++i) hh[i] = dist(rng);
        for (int i = 0; i < nx; ++i) {
            hl[i] = hh[(i - 1 + nx) % nx];
            hr[i] = hh[(i + 1) % nx];
        }
        View1D h("h", nx), grad("grad", nx), L("hl", nx), R("hr", nx), v("v", nx);
        deep_copy(h, hh);
        deep_copy(L, hl);
        deep_copy(R, hr);
        gradient(h, p, dn, grad);
        std::vector<int> idx(nx);
        for (int i = 0; i < nx; ++i) idx[i] = i;
        const real_t dx = Lx / nx;
        rk_site_velocity(idx, nx, dx, p.line_tension, p.driving_force, model, h, L, R, dn, v);
        const std::vector<real_t> g = grad.to_host(), vv = v.to_host();
        double md = 0;
        for (int i = 0; i < nx; ++i) md = std::max(md, std::abs(vv[i] + g[i]));
        CHECK(md < 1e-12);
        // the single-site form with the reference's argument list
        for (int i : {0, 17, nx - 1})
            CHECK(rk_site_velocity(i, nx, dx, p.line_tension, p.driving_force, model, hh[i], hl[i],
                                   hr[i], dn) == vv[i]);
    }
}

// GradientWorkspace: the same gradient, and edge_dvdy is the dy-only sampler
// at each edge midpoint (the wrap edge's at (n - 1) dx / 2)
static void gradient_workspace_edges() {
    const int n = 300;
    const real_t Lx = 3.0;
    FilteredNoise noise(n, 32, Lx, 1.0, 0.5, 0.2, 0.2, 4);
    const DeviceNoise dn = noise.device_noise();
    std::vector<real_t> hh(n);
    for (int i = 0; i < n; ++i) hh[i] = 0.5 + 0.2 * std::sin(0.05 * i);
    View1D h("h", n), g0("g0", n), g1("g1", n);
    deep_copy(h, hh);
    for (const Model model : {Model::Linear, Model::Arclength}) {
        Params p;
        p.physical_size = Lx;
        p.line_tension = 0.8;
        p.model = model;
        GradientWorkspace ws;
        gradient(h, p, dn, g0);
        gradient(h, p, dn, g1, ws);
        CHECK(g0.to_host() == g1.to_host());
        const real_t dx = Lx / n;
        std::vector<real_t> xc(n), hc(n);
        for (int i = 0; i < n; ++i) {
            const int ip = (i + 1) % n;
            xc[i] = (double)(i + ip) * 0.5 * dx;
            hc[i] = (hh[ip] + hh[i]) * 0.5;
        }
        View1D X("x", n), Y("y", n), D("d", n);
        deep_copy(X, xc);
        deep_copy(Y, hc);
        dn.sample_dy(X, Y, D);
        CHECK(ws.edge_dvdy.to_host() == D.to_host());
        if (model == Model::Arclength) {
            const std::vector<real_t> sl = ws.edge_slope.to_host();
            double md = 0;
            for (int i = 0; i < n; ++i) {
                const double s = (hh[(i + 1) % n] - hh[i]) / dx;
                md = std::max(md, std::abs(sl[i] - s / std::sqrt(1 + s * s)));
            }
            CHECK(md < 1e-15);
        }
    }
}

// rk_forward_zero on the washboard of tests/test_rosso_krauth.cpp:24-31 with
// no elasticity: below f_c = A k a site stops where its velocity vanishes;
// above it no zero exists and the march ends at max_advance
static void forward_zero_on_washboard() {
    const SinusoidNoise wash;
    const real_t A = wash.A, k = wash.k;
    RkParams rp;
    rp.dstep = 1.0 / 200;
    rp.max_advance = 1.0;
    rp.root_tol = 1e-12;
    bool found = false;
    const real_t f = 0.5 * A * k;
    const real_t z = rk_forward_zero(3, 16, 1.0, 0.0, f, Model::Linear, 0.3, 0.3, 0.3, wash, rp,
                                     found);
    CHECK(found && z > 0.3);
    CHECK(std::abs(A * k * std::sin(k * (0.3 + z) / 2) + f) < 1e-9);
    const real_t run = rk_forward_zero(3, 16, 1.0, 0.0, 2 * A * k, Model::Linear, 0.3, 0.3, 0.3,
                                       wash, rp, found);
    CHECK(!found && run >= 0.3 + rp.max_advance - 1e-9);
}

int main() {
    initialize();
    reproduces_values_at_nodes();
    derivatives_match_finite_difference();
    periodic_wrap_across_many_cells();
    site_velocity_matches_gradient();
    gradient_workspace_edges();
    forward_zero_on_washboard();
    finalize();
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
