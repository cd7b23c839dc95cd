// qEW CPU ORACLE -- test infrastructure only, never part of the product.
//
// A sequential FP64 restatement of the reference algorithm, used by tests/,
// by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the
// checker (and, built with -fopenmp, as the timed CPU baseline "port").
// Nothing in qew_amd/ links, loads or calls this file.
//
// Every function cites the reference file:line it restates. Summation order is
// the Kokkos-Serial order (i = 0..n-1, sequential). Built with
// -ffp-contract=off so no multiply-add is fused: the product's FP64 GPU kernels
// are compiled the same way and must reproduce the reduction-free quantities
// (gradient, sampler values, Rosso-Krauth configurations) bit for bit.
//
// Pinning status (what checks each part; see tests/test_oracle_*.py):
//   bspline weights / locate / sampler : closed forms, brute-force mod,
//       SciPy map_coordinates(grid-wrap), noise_golden fixture
//   objective / gradient                 : closed-form energies, FD consistency,
//       lbfgs_golden fixture energy
//   FFT / prefilter                      : direct O(n^2) DFT, round trip,
//       SciPy spline_filter(grid-wrap), circulant residual
//   noise generator                      : band limit, amplitude, mt19937_64
//       stream drawn before the mask (statistical + structural checks only)
//   L-BFGS                               : lbfgs_golden e_min/h_min, harmonic
//       and flat minima, branch tests, preconditioner closed form
//   FIRE                                 : flat and harmonic minima, energy
//       decrease, agreement with L-BFGS, the first step against its closed
//       form; iteration counts unpinned (no reference trajectory available)
//   Philox4x32-10                        : published known-answer vectors
//   Langevin                             : equipartition, T=0 descent,
//       force-shifted equilibrium
//   Rosso-Krauth                         : washboard f_c = A k, arrested
//       configuration, velocity == -gradient

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// Potential descriptor: the B-spline table or one of the analytic test noises
// (reference tests/test_qew_model.cpp:22-42, tests/test_lbfgs.cpp:24-38,
// tests/test_dynamics.cpp:20-28, tests/test_rosso_krauth.cpp:24-31).
// ---------------------------------------------------------------------------
// MIDTABLE: coef holds T(ix, iy) = sum_a w_a(1/2) c(ix - 1 + a, iy), the
// x-sum of the bicubic at x-fraction 1/2 (no reference counterpart: the
// device's midpoint-contracted FP32 tables, sampled only at points whose
// x-fraction is 1/2, i.e. at the edge midpoints of lines whose site spacing
// equals the table's; there V = sum_b T(ix, iy_b) wy_b exactly).
enum Kind { TABLE = 0, SINCOS = 1, CONST = 2, ZERO = 3, HARMONIC = 4, WASHBOARD = 5,
            MIDTABLE = 6 };

}  // namespace

extern "C" {

struct orc_noise {
    int kind;
    int nx, ny;
    double dx, dy;
    const double *coef;  // row-major coef[ix*ny + iy] (the reference host layout)
    double p[4];         // analytic parameters
};

struct orc_params {
    double physical_size;
    double line_tension;
    double driving_force;
    int model;  // 0 linear, 1 arclength
};

struct orc_lbfgs_params {
    int m;
    double ftol;
    int max_iter;
    double c1;
    double ls_shrink;
    int max_ls;
    double curvature_eps;
    int precondition;
    double precond_shift;
};

struct orc_lbfgs_result {
    int iterations;
    double max_force;
    int converged;
    long n_func_evals;
    long n_curvature_skips;
    long n_ls_restarts;
};

struct orc_fire_params {
    double dt_max, dt_init, ftol;
    int max_iter, n_min;
    double f_inc, f_dec, alpha_start, f_alpha;
};

struct orc_fire_result {
    int iterations;
    double max_force;
    int converged;
};

struct orc_rk_params {
    double dstep, max_advance, root_tol, sweep_tol, runaway;
    int max_sweeps;
};

struct orc_rk_result {
    int blocked, runaway, sweeps;
    double total_advance, mean_h;
    long n_velocity_evals;  // orc_rk_site_velocity calls (not a reference field)
};

// ---- B-spline basis: src/filtered_noise.h:33-62 ----
void orc_bspline_weights(double t, double *w, double *dw) {
    const double t2 = t * t;
    const double t3 = t2 * t;
    const double omt = 1.0 - t;
    w[0] = omt * omt * omt / 6.0;
    w[1] = (4.0 - 6.0 * t2 + 3.0 * t3) / 6.0;
    w[2] = (1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3) / 6.0;
    w[3] = t3 / 6.0;
    if (dw) {
        dw[0] = -omt * omt / 2.0;
        dw[1] = (-12.0 * t + 9.0 * t2) / 6.0;
        dw[2] = (3.0 + 6.0 * t - 9.0 * t2) / 6.0;
        dw[3] = 3.0 * t2 / 6.0;
    }
}

// ---- DeviceNoise::locate: src/filtered_noise.h:77-101 ----
void orc_locate(const orc_noise *nz, double x, double y, int *ixn, int *iyn,
                double *fx_out, double *fy_out) {
    const int nx = nz->nx, ny = nz->ny;
    double gx = x / nz->dx;
    double gy = y / nz->dy;
    gx -= std::floor(gx / (double)nx) * (double)nx;
    gy -= std::floor(gy / (double)ny) * (double)ny;
    int ix = (int)std::floor(gx);
    int iy = (int)std::floor(gy);
    if (ix >= nx) { ix -= nx; gx -= (double)nx; }
    if (iy >= ny) { iy -= ny; gy -= (double)ny; }
    *fx_out = gx - (double)ix;
    *fy_out = gy - (double)iy;
    ixn[0] = (ix == 0) ? nx - 1 : ix - 1;
    ixn[1] = ix;
    ixn[2] = (ix + 1 == nx) ? 0 : ix + 1;
    ixn[3] = (ixn[2] + 1 == nx) ? 0 : ixn[2] + 1;
    iyn[0] = (iy == 0) ? ny - 1 : iy - 1;
    iyn[1] = iy;
    iyn[2] = (iy + 1 == ny) ? 0 : iy + 1;
    iyn[3] = (iyn[2] + 1 == ny) ? 0 : iyn[2] + 1;
}

// V and dV/dy from a MIDTABLE at (x, y): the cell of orc_locate, the y
// weights, the column's 4 values summed b = 0..3 (dV/dx is not available).
static void orc_mid(const orc_noise *nz, double x, double y, double *v, double *dvdy) {
    int ixn[4], iyn[4];
    double fx, fy;
    orc_locate(nz, x, y, ixn, iyn, &fx, &fy);
    double wy[4], dwy[4];
    orc_bspline_weights(fy, wy, dwy);
    double vv = 0, d = 0;
    for (int b = 0; b < 4; ++b) {
        const double t = nz->coef[(size_t)ixn[1] * nz->ny + iyn[b]];
        vv += t * wy[b];
        d += t * dwy[b];
    }
    *v = vv;
    *dvdy = d / nz->dy;
}

// ---- full sample / value / dy: src/filtered_noise.h:103-176 plus the
//      analytic test functors' sample() ----
void orc_sample(const orc_noise *nz, double x, double y, double *v, double *dvdx,
                double *dvdy) {
    const double *p = nz->p;
    switch (nz->kind) {
    case MIDTABLE:
        orc_mid(nz, x, y, v, dvdy);
        *dvdx = std::nan("");
        return;
    case SINCOS:
        *v = p[0] * std::sin(p[1] * x) * std::cos(p[2] * y);
        *dvdx = p[0] * p[1] * std::cos(p[1] * x) * std::cos(p[2] * y);
        *dvdy = -p[0] * p[2] * std::sin(p[1] * x) * std::sin(p[2] * y);
        return;
    case CONST:
        *v = p[0]; *dvdx = 0; *dvdy = 0;
        return;
    case ZERO:
        *v = 0; *dvdx = 0; *dvdy = 0;
        return;
    case HARMONIC: {
        const double d = y - p[1];
        *v = 0.5 * p[0] * d * d; *dvdx = 0; *dvdy = p[0] * d;
        return;
    }
    case WASHBOARD:
        *v = p[0] * std::cos(p[1] * y); *dvdx = 0;
        *dvdy = -p[0] * p[1] * std::sin(p[1] * y);
        return;
    default: break;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    orc_locate(nz, x, y, ixn, iyn, &fx, &fy);
    double wx[4], dwx[4], wy[4], dwy[4];
    orc_bspline_weights(fx, wx, dwx);
    orc_bspline_weights(fy, wy, dwy);
    double vv = 0, dgx = 0, dgy = 0;
    for (int a = 0; a < 4; ++a) {
        double row = 0, row_dy = 0;
        for (int b = 0; b < 4; ++b) {
            const double c = nz->coef[(size_t)ixn[a] * nz->ny + iyn[b]];
            row += c * wy[b];
            row_dy += c * dwy[b];
        }
        vv += wx[a] * row;
        dgx += dwx[a] * row;
        dgy += wx[a] * row_dy;
    }
    *v = vv;
    *dvdx = dgx / nz->dx;
    *dvdy = dgy / nz->dy;
}

double orc_sample_value(const orc_noise *nz, double x, double y) {
    if (nz->kind == MIDTABLE) {
        double v, d;
        orc_mid(nz, x, y, &v, &d);
        return v;
    }
    if (nz->kind != TABLE) {
        double v, a, b;
        orc_sample(nz, x, y, &v, &a, &b);
        return v;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    orc_locate(nz, x, y, ixn, iyn, &fx, &fy);
    double wx[4], wy[4];
    orc_bspline_weights(fx, wx, nullptr);
    orc_bspline_weights(fy, wy, nullptr);
    double v = 0;
    for (int a = 0; a < 4; ++a) {
        double row = 0;
        for (int b = 0; b < 4; ++b)
            row += nz->coef[(size_t)ixn[a] * nz->ny + iyn[b]] * wy[b];
        v += wx[a] * row;
    }
    return v;
}

double orc_sample_dy(const orc_noise *nz, double x, double y) {
    if (nz->kind == MIDTABLE) {
        double v, d;
        orc_mid(nz, x, y, &v, &d);
        return d;
    }
    if (nz->kind != TABLE) {
        double v, a, b;
        orc_sample(nz, x, y, &v, &a, &b);
        return b;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    orc_locate(nz, x, y, ixn, iyn, &fx, &fy);
    double wx[4], wy[4], dwy[4];
    orc_bspline_weights(fx, wx, nullptr);
    orc_bspline_weights(fy, wy, dwy);
    double d = 0;
    for (int a = 0; a < 4; ++a) {
        double row_dy = 0;
        for (int b = 0; b < 4; ++b)
            row_dy += nz->coef[(size_t)ixn[a] * nz->ny + iyn[b]] * dwy[b];
        d += wx[a] * row_dy;
    }
    return d / nz->dy;
}

}  // extern "C"

namespace {

// ---- objective_eval: src/qew_model.h:53-85 (LineAt / TrialAt :87-99) ----
template <class At>
double objective_eval(int n, const orc_params *p, const orc_noise *nz, const At &at) {
    const double dx = p->physical_size / (double)n;
    const double lt = p->line_tension;
    const double f = p->driving_force;
    double total = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : total) schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const double hi = at(i);
        const double hip = at(ip);
        const double dh_dx = (hip - hi) / dx;
        double line_energy;
        if (p->model == 0)
            line_energy = lt * dh_dx * dh_dx * 0.5;
        else
            line_energy = lt * (std::sqrt(1.0 + dh_dx * dh_dx) - 1.0);
        const double xc = (double)(i + ip) * 0.5 * dx;
        const double hc = (hip + hi) * 0.5;
        total += dx * (line_energy + orc_sample_value(nz, xc, hc) - f * hi);
    }
    return total;
}

}  // namespace

extern "C" {

double orc_objective(int n, const double *h, const orc_params *p, const orc_noise *nz) {
    return objective_eval(n, p, nz, [h](int i) { return h[i]; });
}

double orc_objective_trial(int n, const double *x0, const double *d, double alpha,
                           const orc_params *p, const orc_noise *nz) {
    return objective_eval(n, p, nz, [=](int i) { return x0[i] + alpha * d[i]; });
}

// ---- two-pass gradient: src/qew_model.h:149-194; edvdy / eslope are the
//      GradientWorkspace's edge_dvdy / edge_slope (src/qew_model.h:126-147),
//      eslope written for the arclength model only ----
void orc_gradient_edges(int n, const double *h, const orc_params *p, const orc_noise *nz,
                        double *grad, double *edvdy, double *eslope) {
    const double dx = p->physical_size / (double)n;
    const double lt = p->line_tension;
    const double f = p->driving_force;
    const bool arc = (p->model == 1);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const double xc = (double)(i + ip) * 0.5 * dx;
        const double hc = (h[ip] + h[i]) * 0.5;
        edvdy[i] = orc_sample_dy(nz, xc, hc);
        if (arc) {
            const double dh = (h[ip] - h[i]) / dx;
            eslope[i] = dh / std::sqrt(1.0 + dh * dh);
        }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const int im = (i == 0) ? n - 1 : i - 1;
        double line_force;
        if (!arc) {
            const double d2h = h[im] + h[ip] - 2.0 * h[i];
            line_force = -lt * d2h / (dx * dx);
        } else {
            line_force = lt * (eslope[im] - eslope[i]) / dx;
        }
        grad[i] = dx * (line_force + 0.5 * (edvdy[im] + edvdy[i]) - f);
    }
}

void orc_gradient(int n, const double *h, const orc_params *p, const orc_noise *nz,
                  double *grad) {
    std::vector<double> edvdy(n), eslope(p->model == 1 ? n : 0);
    orc_gradient_edges(n, h, p, nz, grad, edvdy.data(), eslope.data());
}

// Batched noise query used by the sampler tests (reference
// tests/test_filtered_noise.cpp:28-64 sample_points).
void orc_sample_points(const orc_noise *nz, int m, const double *x, const double *y,
                       double *v, double *dvdx, double *dvdy) {
    for (int k = 0; k < m; ++k) orc_sample(nz, x[k], y[k], &v[k], &dvdx[k], &dvdy[k]);
}

}  // extern "C"

// ===========================================================================
// FFT: recursive mixed-radix decimation-in-time, exact twiddles from cos/sin.
// Stands in for kokkos-fft/FFTW at the reference's call sites
// (src/filtered_noise.cpp:58,108,122 and src/lbfgs.h:125-145,175,179).
// sign = -1 forward (unnormalised), +1 backward (caller applies 1/n).
// ===========================================================================
namespace {

int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

// tw[j] = exp(sign * 2 pi i j / N) for the top-level length N; a sub-transform
// of length n uses W_n^e = tw[e * (N / n)].
void fft_rec(int n, const cplx *in, int istride, cplx *out, int sign, const cplx *tw,
             int tws) {
    if (n == 1) { out[0] = in[0]; return; }
    const int r = smallest_factor(n);
    const int m = n / r;
    for (int q = 0; q < r; ++q)
        fft_rec(m, in + (size_t)q * istride, istride * r, out + (size_t)q * m, sign, tw, tws * r);
    std::vector<cplx> t(r);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q)
            t[q] = (q == 0 || k == 0) ? out[(size_t)q * m + k]
                                      : out[(size_t)q * m + k] * tw[(size_t)q * k * tws];
        if (r == 2) {
            out[k] = t[0] + t[1];
            out[(size_t)m + k] = t[0] - t[1];
        } else if (r == 4) {
            // W_4 = exp(sign * i pi / 2) = sign * i
            const cplx a = t[0] + t[2], b = t[0] - t[2];
            const cplx c = t[1] + t[3], e = t[1] - t[3];
            const cplx je = (sign < 0) ? cplx(e.imag(), -e.real()) : cplx(-e.imag(), e.real());
            out[k] = a + c;
            out[(size_t)m + k] = b + je;
            out[(size_t)2 * m + k] = a - c;
            out[(size_t)3 * m + k] = b - je;
        } else {
            for (int s = 0; s < r; ++s) {
                cplx acc = t[0];
                for (int q = 1; q < r; ++q)
                    acc += t[q] * tw[(size_t)(((long long)q * s) % r) * m * tws];
                out[(size_t)s * m + k] = acc;
            }
        }
    }
}

std::vector<cplx> twiddles(int n, int sign) {
    std::vector<cplx> tw(n);
    for (int j = 0; j < n; ++j) {
        const double ang = (double)sign * 2.0 * M_PI * (double)j / (double)n;
        tw[j] = cplx(std::cos(ang), std::sin(ang));
    }
    return tw;
}

void fft1(int n, const cplx *in, cplx *out, int sign) {
    const std::vector<cplx> tw = twiddles(n, sign);
    fft_rec(n, in, 1, out, sign, tw.data(), 1);
}

}  // namespace

extern "C" {

// In-place complex FFT of length n on interleaved (re, im) doubles.
void orc_fft(int n, double *data, int sign) {
    std::vector<cplx> in(n), out(n);
    for (int i = 0; i < n; ++i) in[i] = cplx(data[2 * i], data[2 * i + 1]);
    fft1(n, in.data(), out.data(), sign);
    for (int i = 0; i < n; ++i) { data[2 * i] = out[i].real(); data[2 * i + 1] = out[i].imag(); }
}

// 2-D FFT of an nx x ny row-major complex array (index i*ny + j); backward
// applies 1/(nx*ny) like numpy ifft2 / kokkos-fft backward normalisation.
void orc_fft2(int nx, int ny, double *data, int sign) {
    const std::vector<cplx> twy = twiddles(ny, sign), twx = twiddles(nx, sign);
    std::vector<cplx> a(ny), b(ny);
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) a[j] = cplx(data[2 * ((size_t)i * ny + j)], data[2 * ((size_t)i * ny + j) + 1]);
        fft_rec(ny, a.data(), 1, b.data(), sign, twy.data(), 1);
        for (int j = 0; j < ny; ++j) { data[2 * ((size_t)i * ny + j)] = b[j].real(); data[2 * ((size_t)i * ny + j) + 1] = b[j].imag(); }
    }
    std::vector<cplx> c(nx), d(nx);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) c[i] = cplx(data[2 * ((size_t)i * ny + j)], data[2 * ((size_t)i * ny + j) + 1]);
        fft_rec(nx, c.data(), 1, d.data(), sign, twx.data(), 1);
        for (int i = 0; i < nx; ++i) { data[2 * ((size_t)i * ny + j)] = d[i].real(); data[2 * ((size_t)i * ny + j) + 1] = d[i].imag(); }
    }
    if (sign > 0) {
        const double s = 1.0 / ((double)nx * (double)ny);
        for (size_t k = 0; k < 2 * (size_t)nx * ny; ++k) data[k] *= s;
    }
}

// ---- periodic cubic B-spline prefilter: src/filtered_noise.cpp:92-136 ----
void orc_prefilter(int nx, int ny, const double *field, double *coef) {
    const double two_pi = 2.0 * std::acos(-1.0);
    std::vector<double> S(2 * (size_t)nx * ny);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) { S[2 * k] = field[k]; S[2 * k + 1] = 0; }
    orc_fft2(nx, ny, S.data(), -1);
    for (int i = 0; i < nx; ++i) {
        const double kx = 2.0 / 3 + std::cos(two_pi * i / nx) / 3.0;
        for (int j = 0; j < ny; ++j) {
            const double ky = 2.0 / 3 + std::cos(two_pi * j / ny) / 3.0;
            const double den = kx * ky;
            const size_t k = (size_t)i * ny + j;
            S[2 * k] /= den;
            S[2 * k + 1] /= den;
        }
    }
    orc_fft2(nx, ny, S.data(), +1);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) coef[k] = S[2 * k];
}

// ---- FilteredNoise(nx,ny,Lx,Ly,amp,xi_x,xi_y,seed): src/filtered_noise.cpp:25-82.
// Writes field (noise_on_grid) and coef, both row-major nx*ny. ----
void orc_filtered_noise(int nx, int ny, double Lx, double Ly, double amplitude,
                        double xi_x, double xi_y, uint64_t seed, double *field,
                        double *coef) {
    const double dx = Lx / nx, dy = Ly / ny;
    const double two_pi = 2.0 * std::acos(-1.0);
    std::vector<double> A(2 * (size_t)nx * ny);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> phase(0, two_pi);
    auto fftfreq = [](int k, int n) {
        const int ks = (k < (n + 1) / 2) ? k : k - n;
        return (double)ks / (double)n;
    };
    for (int i = 0; i < nx; ++i) {
        const double fx = fftfreq(i, nx) / dx;
        for (int j = 0; j < ny; ++j) {
            const double fy = fftfreq(j, ny) / dy;
            const double ph = phase(rng);  // drawn before the mask test
            const size_t k = (size_t)i * ny + j;
            if ((fx * xi_x) * (fx * xi_x) + (fy * xi_y) * (fy * xi_y) > 1.0) {
                A[2 * k] = 0; A[2 * k + 1] = 0;
            } else {
                A[2 * k] = std::cos(ph); A[2 * k + 1] = std::sin(ph);
            }
        }
    }
    orc_fft2(nx, ny, A.data(), +1);
    const size_t n = (size_t)nx * ny;
    double mean = 0;
    for (size_t k = 0; k < n; ++k) mean += A[2 * k];
    mean /= (double)n;
    double var = 0;
    for (size_t k = 0; k < n; ++k) { const double d = A[2 * k] - mean; var += d * d; }
    const double scale = amplitude / std::sqrt(var / (double)n);
    for (size_t k = 0; k < n; ++k) field[k] = A[2 * k] * scale;
    orc_prefilter(nx, ny, field, coef);
}

}  // extern "C"

// ===========================================================================
// L-BFGS: src/lbfgs.h:94-287
// ===========================================================================
namespace {

double dot(int n, const double *a, const double *b) {
    double s = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

double max_abs(int n, const double *a) {
    double s = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(max : s) schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const double v = std::fabs(a[i]);
        if (v > s) s = v;
    }
    return s;
}

// d = irfft(rfft(q) / eig) with the backward 1/n normalisation
// (src/lbfgs.h:173-179). The half spectrum k = 0..n/2 is what rfft keeps.
void precondition_apply(int n, const double *q, const std::vector<double> &eig, double *d) {
    std::vector<double> buf(2 * (size_t)n);
    for (int i = 0; i < n; ++i) { buf[2 * i] = q[i]; buf[2 * i + 1] = 0; }
    orc_fft(n, buf.data(), -1);
    const int nh = n / 2 + 1;
    std::vector<cplx> half(nh);
    for (int k = 0; k < nh; ++k) half[k] = cplx(buf[2 * k], buf[2 * k + 1]) / eig[k];
    // irfft: Hermitian extension; the imaginary parts of the DC and (even n)
    // Nyquist bins are ignored, as a c2r transform does.
    for (int k = 0; k < n; ++k) {
        cplx v;
        if (k < nh) v = half[k];
        else v = std::conj(half[n - k]);
        if (k == 0 || (n % 2 == 0 && k == n / 2)) v = cplx(v.real(), 0);
        buf[2 * k] = v.real(); buf[2 * k + 1] = v.imag();
    }
    orc_fft(n, buf.data(), +1);
    for (int i = 0; i < n; ++i) d[i] = buf[2 * i] / (double)n;
}

}  // namespace

extern "C" {

// Preconditioner eigenvalues eig[k], k = 0..n/2 (src/lbfgs.h:127-143).
void orc_precond_eig(int n, double kappa, double physical_size, double precond_shift,
                     double *eig) {
    const double dx = physical_size / (double)n;
    const double pi = 3.14159265358979323846;
    double shift = precond_shift;
    if (shift <= 0) {
        const double s1 = std::sin(pi / (double)n);
        shift = (kappa / dx) * 4 * s1 * s1;
    }
    for (int k = 0; k < n / 2 + 1; ++k) {
        const double s = std::sin(pi * (double)k / (double)n);
        eig[k] = (kappa / dx) * 4 * s * s + shift;
    }
}

void orc_precondition(int n, const double *q, const double *eig, double *d) {
    std::vector<double> e(eig, eig + n / 2 + 1);
    precondition_apply(n, q, e, d);
}

void orc_lbfgs(int n, double *h, const orc_params *p, const orc_noise *nz,
               const orc_lbfgs_params *lp, orc_lbfgs_result *res_out) {
    const int m = lp->m;
    std::vector<double> g(n), g_new(n), d(n), q(n), s_tmp(n), y_tmp(n);
    std::vector<std::vector<double>> S(m, std::vector<double>(n)), Y(m, std::vector<double>(n));
    std::vector<double> rho(m, 0.0);
    std::vector<int> slots;
    double gamma = 1.0;
    std::vector<double> eig;
    if (lp->precondition) {
        eig.resize(n / 2 + 1);
        orc_precond_eig(n, p->line_tension, p->physical_size, lp->precond_shift, eig.data());
    }
    orc_gradient(n, h, p, nz, g.data());
    orc_lbfgs_result res{};
    int ls_fail_streak = 0;
    double phi = 0;
    for (int it = 0; it < lp->max_iter; ++it) {
        const double gmax = max_abs(n, g.data());
        if (gmax < lp->ftol) {
            res.iterations = it; res.max_force = gmax; res.converged = 1;
            *res_out = res;
            return;
        }
        q = g;
        const int hs = (int)slots.size();
        std::vector<double> la(hs);
        for (int idx = hs - 1; idx >= 0; --idx) {
            const int row = slots[idx];
            const double a = rho[row] * dot(n, S[row].data(), q.data());
            la[idx] = a;
            const double c = -a;
            for (int i = 0; i < n; ++i) q[i] += c * Y[row][i];
        }
        if (lp->precondition) {
            precondition_apply(n, q.data(), eig, d.data());
        } else {
            const double g0 = slots.empty() ? 1.0 : gamma;
            for (int i = 0; i < n; ++i) d[i] = g0 * q[i];
        }
        for (int idx = 0; idx < hs; ++idx) {
            const int row = slots[idx];
            const double b = rho[row] * dot(n, Y[row].data(), d.data());
            const double c = la[idx] - b;
            for (int i = 0; i < n; ++i) d[i] += c * S[row][i];
        }
        for (int i = 0; i < n; ++i) d[i] = -d[i];
        double dphi0 = dot(n, g.data(), d.data());
        if (dphi0 >= 0) {
            slots.clear();
            for (int i = 0; i < n; ++i) d[i] = -g[i];
            dphi0 = -dot(n, g.data(), g.data());
        }
        if (it == 0) {
            phi = orc_objective(n, h, p, nz);
            ++res.n_func_evals;
        }
        const double phi0 = phi;
        double alpha = 1.0;
        if (!lp->precondition && slots.empty())
            alpha = 1.0 / std::max(1.0, std::sqrt(dot(n, g.data(), g.data())));
        bool ok = false;
        for (int ls = 0; ls < lp->max_ls; ++ls) {
            const double phi_trial = orc_objective_trial(n, h, d.data(), alpha, p, nz);
            ++res.n_func_evals;
            if (phi_trial <= phi0 + lp->c1 * alpha * dphi0) {
                phi = phi_trial;
                ok = true;
                break;
            }
            alpha *= lp->ls_shrink;
        }
        if (!ok) {
            if (slots.empty() || ++ls_fail_streak > 2) {
                res.iterations = it; res.max_force = gmax; res.converged = 0;
                *res_out = res;
                return;
            }
            slots.clear();
            ++res.n_ls_restarts;
            continue;
        }
        ls_fail_streak = 0;
        const double al = alpha;
        for (int i = 0; i < n; ++i) h[i] += al * d[i];
        orc_gradient(n, h, p, nz, g_new.data());
        double sy = 0, ss = 0, yy = 0;
        for (int i = 0; i < n; ++i) {
            const double s = al * d[i];
            const double y = g_new[i] - g[i];
            s_tmp[i] = s;
            y_tmp[i] = y;
            sy += s * y;
            ss += s * s;
            yy += y * y;
        }
        if (sy > lp->curvature_eps * std::sqrt(ss * yy)) {
            int row;
            if ((int)slots.size() < m) {
                row = (int)slots.size();
            } else {
                row = slots.front();
                slots.erase(slots.begin());
            }
            std::swap(S[row], s_tmp);
            std::swap(Y[row], y_tmp);
            rho[row] = 1.0 / sy;
            gamma = sy / yy;
            slots.push_back(row);
        } else {
            ++res.n_curvature_skips;
        }
        std::swap(g, g_new);
    }
    res.iterations = lp->max_iter;
    res.max_force = max_abs(n, g.data());
    res.converged = 0;
    *res_out = res;
}

// FIRE, semi-implicit Euler variant (src/fire.h:41-131): v += dt F, the power
// P = F.v adapts dt and alpha (and zeroes v when P <= 0), v is mixed toward
// F with weight alpha |v| / |F|, then h += dt v; F = -grad. Sums are
// sequential in i.
void orc_fire(int n, double *h, const orc_params *p, const orc_noise *nz,
              const orc_fire_params *fp, orc_fire_result *res_out) {
    std::vector<double> g(n), v(n, 0.0);
    double dt = fp->dt_init > 0 ? fp->dt_init : 0.1 * fp->dt_max;
    double alpha = fp->alpha_start;
    int n_pos = 0;
    orc_gradient(n, h, p, nz, g.data());
    orc_fire_result res{};
    for (int it = 0; it < fp->max_iter; ++it) {
        const double max_f = max_abs(n, g.data());
        if (max_f < fp->ftol) {
            res.iterations = it; res.max_force = max_f; res.converged = 1;
            *res_out = res;
            return;
        }
        for (int i = 0; i < n; ++i) v[i] -= dt * g[i];
        double P = 0;
        for (int i = 0; i < n; ++i) P += -g[i] * v[i];
        if (P > 0) {
            if (n_pos > fp->n_min) {
                dt = std::fmin(dt * fp->f_inc, fp->dt_max);
                alpha *= fp->f_alpha;
            }
            ++n_pos;
        } else {
            for (int i = 0; i < n; ++i) v[i] = 0.0;
            dt *= fp->f_dec;
            alpha = fp->alpha_start;
            n_pos = 0;
        }
        double nv2 = 0, nf2 = 0;
        for (int i = 0; i < n; ++i) nv2 += v[i] * v[i];
        for (int i = 0; i < n; ++i) nf2 += g[i] * g[i];
        const double nf = std::sqrt(nf2);
        if (nf > 0) {
            const double a = alpha;
            const double scale = alpha * std::sqrt(nv2) / nf;
            for (int i = 0; i < n; ++i) v[i] = (1.0 - a) * v[i] - scale * g[i];
        }
        for (int i = 0; i < n; ++i) h[i] += dt * v[i];
        orc_gradient(n, h, p, nz, g.data());
    }
    res.iterations = fp->max_iter;
    res.max_force = max_abs(n, g.data());
    res.converged = 0;
    *res_out = res;
}

// ===========================================================================
// Philox4x32-10 (Salmon et al., SC'11 "Parallel random numbers: as easy as
// 1, 2, 3") and the counter-based normal used by the Langevin step. The
// reference draws from a Kokkos XorShift64 pool (src/dynamics.h:29,49-58),
// whose stream is scheduling-dependent; the GPU engine keys every normal by
// (seed, step, site) instead, and so does this restatement.
// ===========================================================================
void orc_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key_in, uint32_t *out) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Standard normal for (seed, step, site): Philox4x32-10 with key = seed and
// counter = (site/2, step_lo, step_hi, 0); two 53-bit uniforms u1 in (0,1],
// u2 in [0,1); Box-Muller radius sqrt(-2 ln u1) times cos (even site) or sin
// (odd site) of 2 pi u2.
double orc_normal(uint64_t seed, uint64_t step, uint64_t site) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)(site >> 1), (uint32_t)step, (uint32_t)(step >> 32), 0u};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    const uint64_t k1 = ((((uint64_t)r[0]) << 32) | r[1]) >> 11;
    const uint64_t k2 = ((((uint64_t)r[2]) << 32) | r[3]) >> 11;
    const double u1 = (double)(k1 + 1) * 0x1.0p-53;
    const double u2 = (double)k2 * 0x1.0p-53;
    const double rad = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586 * u2;
    return (site & 1) ? rad * std::sin(th) : rad * std::cos(th);
}

// ---- LangevinDynamics::step: src/dynamics.h:34-64 with the Philox normal.
//      `step` is the integrator's step counter (0 for the first call). ----
void orc_langevin_step(int n, double *h, const orc_params *p, const orc_noise *nz,
                       double a, double b, uint64_t seed, uint64_t step) {
    std::vector<double> g(n);
    orc_gradient(n, h, p, nz, g.data());
    if (b > 0) {
        for (int i = 0; i < n; ++i) h[i] += -a * g[i] + b * orc_normal(seed, step, (uint64_t)i);
    } else {
        for (int i = 0; i < n; ++i) h[i] += -a * g[i];
    }
}

// ===========================================================================
// Rosso-Krauth: src/rosso_krauth.h:52-194
// ===========================================================================
double orc_rk_site_velocity(int i, int n, double dx, double lt, double f, int model,
                            double hi, double hl, double hr, const orc_noise *nz) {
    const int im = (i == 0) ? n - 1 : i - 1;
    const int ip = (i + 1 == n) ? 0 : i + 1;
    double line_force;
    if (model == 0) {
        line_force = -lt * (hl + hr - 2.0 * hi) / (dx * dx);
    } else {
        const double dhr = (hr - hi) / dx;
        const double dhl = (hi - hl) / dx;
        line_force = lt * (dhl / std::sqrt(1.0 + dhl * dhl) - dhr / std::sqrt(1.0 + dhr * dhr)) / dx;
    }
    const double xcl = (double)(im + i) * 0.5 * dx;
    const double hcl = (hl + hi) * 0.5;
    const double xcr = (double)(i + ip) * 0.5 * dx;
    const double hcr = (hi + hr) * 0.5;
    const double dvl = orc_sample_dy(nz, xcl, hcl);
    const double dvr = orc_sample_dy(nz, xcr, hcr);
    const double grad = dx * (line_force + 0.5 * (dvl + dvr) - f);
    return -grad;
}

double orc_rk_forward_zero(int i, int n, double dx, double lt, double f, int model,
                           double hi, double hl, double hr, const orc_noise *nz,
                           const orc_rk_params *rp, int *found, long *evals) {
    const double v0 = orc_rk_site_velocity(i, n, dx, lt, f, model, hi, hl, hr, nz);
    *evals = 1;
    if (v0 <= 0) { *found = 1; return hi; }
    double a = hi;
    double advanced = 0;
    while (advanced < rp->max_advance) {
        const double b = a + rp->dstep;
        const double vb = orc_rk_site_velocity(i, n, dx, lt, f, model, b, hl, hr, nz);
        ++*evals;
        if (vb <= 0) {
            double lo = a, hicur = b;
            for (int k = 0; k < 60 && (hicur - lo) > rp->root_tol; ++k) {
                const double mid = 0.5 * (lo + hicur);
                const double vm = orc_rk_site_velocity(i, n, dx, lt, f, model, mid, hl, hr, nz);
                ++*evals;
                if (vm > 0) lo = mid; else hicur = mid;
            }
            *found = 1;
            return 0.5 * (lo + hicur);
        }
        a = b;
        advanced += rp->dstep;
    }
    *found = 0;
    return a;
}

// Returns 0, or -1 for odd n (the reference throws std::invalid_argument).
int orc_rk_block(int n, double *h, const orc_params *p, const orc_noise *nz,
                 const orc_rk_params *rp, orc_rk_result *out) {
    if (n % 2 != 0) return -1;
    const double dx = p->physical_size / (double)n;
    const double lt = p->line_tension;
    const double f = p->driving_force;
    std::vector<double> adv(n);
    std::vector<long> evals(n);
    double mean_h0 = 0;
    for (int i = 0; i < n; ++i) mean_h0 += h[i];
    mean_h0 /= (double)n;
    orc_rk_result res{};
    double cum_adv = 0;
    for (int sweep = 0; sweep < rp->max_sweeps; ++sweep) {
        for (int color = 0; color < 2; ++color) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
            for (int i = color; i < n; i += 2) {
                const int im = (i == 0) ? n - 1 : i - 1;
                const int ip = (i + 1 == n) ? 0 : i + 1;
                int found = 0;
                const double hnew = orc_rk_forward_zero(i, n, dx, lt, f, p->model, h[i], h[im], h[ip], nz, rp, &found, &evals[i]);
                adv[i] = hnew - h[i];
                h[i] = hnew;
            }
        }
        double total_adv = 0;
        for (int i = 0; i < n; ++i) total_adv += adv[i];
        for (int i = 0; i < n; ++i) res.n_velocity_evals += evals[i];
        cum_adv += total_adv;
        const double mean_adv = cum_adv / (double)n;
        res.sweeps = sweep + 1;
        res.total_advance = total_adv;
        res.mean_h = mean_h0 + mean_adv;
        if (mean_adv > rp->runaway) { res.runaway = 1; res.blocked = 0; *out = res; return 0; }
        if (total_adv < rp->sweep_tol) { res.blocked = 1; res.runaway = 0; *out = res; return 0; }
    }
    res.blocked = 0; res.runaway = 0;
    *out = res;
    return 0;
}

// Sets the OpenMP thread count of later calls (no effect on the sequential
// oracle).
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Number of OpenMP threads this library uses (1 for the sequential oracle).
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

// ===========================================================================
// Drivers over many steps / many lines (the reference's drivers loop over
// these calls one by one: src/qew_runs.h:152-175 for the anneal, :70-105 for
// the static sweep). Same arithmetic as the single calls above; in the
// OpenMP build the work is spread over threads without changing any result
// (the kicks are keyed by (seed, step, site), the gradient is reduction-free,
// and each line's L-BFGS runs on one thread).
// ===========================================================================
extern "C" {

// `nsteps` Langevin steps with per-step coefficients a[k], b[k] and
// integrator step numbers step0 + k (src/dynamics.h:34-64 per step).
void orc_langevin_run(int n, double *h, const orc_params *p, const orc_noise *nz, int nsteps,
                      const double *a, const double *b, uint64_t seed, uint64_t step0) {
    std::vector<double> g(n);
    for (int k = 0; k < nsteps; ++k) {
        orc_gradient(n, h, p, nz, g.data());
        const double ak = a[k], bk = b[k];
        const uint64_t step = step0 + (uint64_t)k;
        if (bk > 0) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
            for (int i = 0; i < n; ++i) h[i] += -ak * g[i] + bk * orc_normal(seed, step, (uint64_t)i);
        } else {
            for (int i = 0; i < n; ++i) h[i] += -ak * g[i];
        }
    }
}

// orc_lbfgs on every line of a ragged batch: line l is h[off[l] .. off[l+1])
// with parameters p[l], potential nz[l] and preconditioner shift shift[l]
// (> 0 replaces lp->precond_shift). Lines are spread over OpenMP threads
// (dynamic schedule, longest first as listed by `order`, or in index order
// when order is null); each line runs sequentially on one thread.
void orc_lbfgs_batch(int nlines, const int64_t *off, double *h, const orc_params *p,
                     const orc_noise *nz, const orc_lbfgs_params *lp, const double *shift,
                     const int *order, orc_lbfgs_result *res) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int k = 0; k < nlines; ++k) {
        const int l = order ? order[k] : k;
        orc_lbfgs_params lpl = *lp;
        if (shift && shift[l] > 0) lpl.precond_shift = shift[l];
        orc_lbfgs((int)(off[l + 1] - off[l]), h + off[l], &p[l], &nz[l], &lpl, &res[l]);
    }
}

}  // extern "C"
