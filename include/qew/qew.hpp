// C++ facade of the MI355X qEW engine: the reference's API names and
// semantics (IMTEK-Simulation/qEW src/*.h) on top of the C ABI in qew_amd.h.
// Header-only; link against libqew_amd.so.
//
//   reference                          here
//   Kokkos::initialize/finalize        qew::ScopeGuard / initialize / finalize
//   View1D (src/qew_types.h:17)        qew::View1D (device FP64 vector)
//   FilteredNoise, DeviceNoise         qew::FilteredNoise, qew::DeviceNoise
//   Noise functors of the tests        qew::ZeroNoise, ConstNoise, HarmonicWell,
//                                      SinusoidNoise, AnalyticNoise
//   objective / gradient               qew::objective / qew::gradient
//   lbfgs_minimize                     qew::lbfgs_minimize
//   fire_minimize                      qew::fire_minimize
//   LangevinDynamics                   qew::LangevinDynamics
//   NoiseSample (src/qew_types.h:26)   qew::NoiseSample
//   rk_site_velocity, rk_forward_zero,
//   rk_block                           qew::rk_site_velocity, rk_forward_zero,
//                                      rk_block
//   static_sweep / dynamic_anneal /
//   depinning_ramp                     qew::static_sweep / dynamic_anneal /
//                                      depinning_ramp (qew_runs.h)
//
// Errors: QEW_ERR_INVALID_ARGUMENT becomes std::invalid_argument (the
// reference throws it for odd line lengths in rk_block), every other failure
// std::runtime_error. Non-convergence is reported in the result structs.
#ifndef QEW_FACADE_HPP
#define QEW_FACADE_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../qew_amd.h"

namespace qew {

using real_t = double;

inline void check(qew_status s) {
    if (s == QEW_OK) return;
    const std::string msg = qew_last_error();
    if (s == QEW_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ---- process-wide context (the default execution space instance) --------
namespace detail {
inline qew_ctx &ctx_slot() {
    static qew_ctx c = nullptr;
    return c;
}
}  // namespace detail

inline void initialize(int device = 0) {
    if (!detail::ctx_slot()) check(qew_ctx_create(device, nullptr, &detail::ctx_slot()));
}

// Drops --kokkos-* options from argv (the reference's mains call
// Kokkos::initialize(argc, argv), which consumes them; README.md:145-147).
inline void initialize(int &argc, char **argv, int device = 0) {
    int w = 1;
    for (int r = 1; r < argc; ++r) {
        const std::string a = argv[r];
        if (a.rfind("--kokkos-", 0) == 0) {
            if (a.find('=') == std::string::npos && r + 1 < argc &&
                std::string(argv[r + 1]).rfind("--", 0) != 0)
                ++r;  // its value
            continue;
        }
        argv[w++] = argv[r];
    }
    argc = w;
    initialize(device);
}

inline void finalize() {
    if (detail::ctx_slot()) {
        qew_ctx_destroy(detail::ctx_slot());
        detail::ctx_slot() = nullptr;
    }
}

inline qew_ctx context() {
    initialize();
    return detail::ctx_slot();
}

struct ScopeGuard {
    ScopeGuard(int &argc, char **argv) { initialize(argc, argv); }
    ~ScopeGuard() { finalize(); }
};

inline void fence() { check(qew_ctx_synchronize(context())); }

// ---- View1D: a device vector of doubles ----------------------------------
class View1D {
public:
    View1D() = default;
    explicit View1D(const std::string &label, size_t n) : label_(label) { allocate(n); }
    View1D(size_t n) { allocate(n); }

    size_t extent(int) const { return n_; }
    size_t size() const { return n_; }
    double *data() const { return ptr_.get(); }
    const std::string &label() const { return label_; }

    void copy_from_host(const double *src) {
        if (n_) check(qew_copy_to_device(context(), data(), src, sizeof(double) * n_));
    }
    void copy_to_host(double *dst) const {
        if (n_) check(qew_copy_to_host(context(), dst, data(), sizeof(double) * n_));
    }
    std::vector<double> to_host() const {
        std::vector<double> v(n_);
        copy_to_host(v.data());
        return v;
    }
    void fill(double value) {
        std::vector<double> v(n_, value);
        copy_from_host(v.data());
    }

private:
    void allocate(size_t n) {
        n_ = n;
        void *p = nullptr;
        check(qew_device_alloc(context(), sizeof(double) * (n ? n : 1), &p));
        qew_ctx c = context();
        ptr_ = std::shared_ptr<double>(static_cast<double *>(p),
                                       [c](double *q) { qew_device_free(c, q); });
    }
    std::shared_ptr<double> ptr_;
    size_t n_ = 0;
    std::string label_;
};

// A device array of ints (site indices of the per-site Rosso-Krauth calls).
class DeviceInts {
public:
    explicit DeviceInts(const std::vector<int> &v) : n_(v.size()) {
        void *p = nullptr;
        check(qew_device_alloc(context(), sizeof(int) * (n_ ? n_ : 1), &p));
        qew_ctx c = context();
        ptr_ = std::shared_ptr<int>(static_cast<int *>(p), [c](int *q) { qew_device_free(c, q); });
        if (n_) check(qew_copy_to_device(context(), ptr_.get(), v.data(), sizeof(int) * n_));
    }
    const int *data() const { return ptr_.get(); }
    size_t size() const { return n_; }

private:
    std::shared_ptr<int> ptr_;
    size_t n_;
};

inline void deep_copy(const View1D &dst, double value) { const_cast<View1D &>(dst).fill(value); }
inline void deep_copy(const View1D &dst, const std::vector<double> &src) {
    const_cast<View1D &>(dst).copy_from_host(src.data());
}
inline std::vector<double> to_host_vec(const View1D &h) { return h.to_host(); }

// ---- potentials ----------------------------------------------------------
// Value and partial derivatives of the pinning potential at a point.
struct NoiseSample {
    real_t v;
    real_t dv_dx;
    real_t dv_dy;
};

// A potential on the device (the reference's Noise concept,
// src/qew_model.h:31-48). Shared handle: copies are cheap, like DeviceNoise.
// The reference's sample / sample_value / sample_dy are device functions
// called inside its kernels; here they evaluate device views of points in
// one launch, or host points through one launch per call.
class DeviceNoise {
public:
    DeviceNoise() = default;
    explicit DeviceNoise(qew_noise h) : h_(h, [](qew_noise_s *p) { qew_noise_destroy(p); }) {}
    qew_noise handle() const { return h_.get(); }
    int nx() const { return info().nx; }
    int ny() const { return info().ny; }

    // src/filtered_noise.h:103-134: value and both derivatives at (x[k], y[k])
    void sample(const View1D &x, const View1D &y, const View1D &v, const View1D &dv_dx,
                const View1D &dv_dy) const {
        const size_t m = same_extent(x, y, v);
        if (dv_dx.extent(0) != m || dv_dy.extent(0) != m)
            throw std::invalid_argument("sample: views of different extents");
        check(qew_noise_sample(context(), handle(), (int)m, x.data(), y.data(), v.data(),
                               dv_dx.data(), dv_dy.data()));
    }
    // src/filtered_noise.h:138-155: the value-only sampler (objective)
    void sample_value(const View1D &x, const View1D &y, const View1D &v) const {
        const size_t m = same_extent(x, y, v);
        check(qew_noise_sample_value(context(), handle(), (int)m, x.data(), y.data(), v.data()));
    }
    // src/filtered_noise.h:159-176: the dy-only sampler (gradient, site velocity)
    void sample_dy(const View1D &x, const View1D &y, const View1D &dv_dy) const {
        const size_t m = same_extent(x, y, dv_dy);
        check(qew_noise_sample_dy(context(), handle(), (int)m, x.data(), y.data(), dv_dy.data()));
    }
    std::vector<NoiseSample> sample(const std::vector<real_t> &x,
                                    const std::vector<real_t> &y) const {
        if (x.size() != y.size()) throw std::invalid_argument("sample: x and y differ in length");
        const size_t m = x.size();
        View1D X(m), Y(m), V(m), DX(m), DY(m);
        X.copy_from_host(x.data());
        Y.copy_from_host(y.data());
        sample(X, Y, V, DX, DY);
        const std::vector<double> v = V.to_host(), dx = DX.to_host(), dy = DY.to_host();
        std::vector<NoiseSample> out(m);
        for (size_t k = 0; k < m; ++k) out[k] = NoiseSample{v[k], dx[k], dy[k]};
        return out;
    }
    NoiseSample sample(real_t x, real_t y) const { return sample(std::vector<real_t>{x}, {y})[0]; }
    real_t sample_value(real_t x, real_t y) const { return point(x, y, false); }
    real_t sample_dy(real_t x, real_t y) const { return point(x, y, true); }

private:
    struct Info {
        int kind, dtype, nx, ny;
        double dx, dy;
    };
    Info info() const {
        Info i{};
        check(qew_noise_info(handle(), &i.kind, &i.dtype, &i.nx, &i.ny, &i.dx, &i.dy));
        return i;
    }
    static size_t same_extent(const View1D &a, const View1D &b, const View1D &c) {
        if (a.extent(0) != b.extent(0) || a.extent(0) != c.extent(0))
            throw std::invalid_argument("sample: views of different extents");
        return a.extent(0);
    }
    real_t point(real_t x, real_t y, bool dy) const {
        View1D X(1), Y(1), R(1);
        X.copy_from_host(&x);
        Y.copy_from_host(&y);
        if (dy) sample_dy(X, Y, R);
        else sample_value(X, Y, R);
        return R.to_host()[0];
    }
    std::shared_ptr<qew_noise_s> h_;
};

inline DeviceNoise make_analytic(int kind, double a = 0, double b = 0, double c = 0) {
    const double p[4] = {a, b, c, 0};
    qew_noise h = nullptr;
    check(qew_noise_analytic(context(), kind, p, &h));
    return DeviceNoise(h);
}

// tests/test_qew_model.cpp:22-42, tests/test_lbfgs.cpp:24-38,
// tests/test_rosso_krauth.cpp:24-31
struct AnalyticNoise {
    real_t amp = 0.3, kx = 0.7, ky = 1.3;
    DeviceNoise device() const { return make_analytic(QEW_NOISE_SINCOS, amp, kx, ky); }
};
struct ConstNoise {
    real_t v0 = 0;
    DeviceNoise device() const { return make_analytic(QEW_NOISE_CONST, v0); }
};
struct ZeroNoise {
    DeviceNoise device() const { return make_analytic(QEW_NOISE_ZERO); }
};
struct HarmonicWell {
    real_t k = 2.0, y0 = 0.5;
    DeviceNoise device() const { return make_analytic(QEW_NOISE_HARMONIC, k, y0); }
};
struct SinusoidNoise {
    real_t A = 0.1, k = 2.0 * 3.14159265358979323846;
    DeviceNoise device() const { return make_analytic(QEW_NOISE_WASHBOARD, A, k); }
};

inline DeviceNoise as_device(const DeviceNoise &n) { return n; }
template <class N>
DeviceNoise as_device(const N &n) {
    return n.device();
}

// FilteredNoise (src/filtered_noise.h:183-244): generated on the device.
class FilteredNoise {
public:
    FilteredNoise(int nx, int ny, real_t Lx, real_t Ly, real_t amplitude, real_t xi_x, real_t xi_y,
                  std::uint64_t seed, int dtype = QEW_F64)
        : nx_(nx), ny_(ny), Lx_(Lx), Ly_(Ly) {
        qew_noise h = nullptr;
        check(qew_noise_generate(context(), nx, ny, Lx, Ly, amplitude, xi_x, xi_y, seed, dtype, &h));
        dn_ = DeviceNoise(h);
    }
    FilteredNoise(const real_t *field, int nx, int ny, real_t Lx, real_t Ly, int dtype = QEW_F64)
        : nx_(nx), ny_(ny), Lx_(Lx), Ly_(Ly) {
        qew_noise h = nullptr;
        check(qew_noise_from_field(context(), field, nx, ny, Lx, Ly, dtype, &h));
        dn_ = DeviceNoise(h);
    }
    DeviceNoise device_noise() const { return dn_; }
    std::vector<real_t> field() const {
        std::vector<real_t> f((size_t)nx_ * ny_);
        check(qew_noise_field(dn_.handle(), f.data()));
        return f;
    }
    int nx() const { return nx_; }
    int ny() const { return ny_; }
    real_t dx() const { return Lx_ / nx_; }
    real_t dy() const { return Ly_ / ny_; }

private:
    int nx_, ny_;
    real_t Lx_, Ly_;
    DeviceNoise dn_;
};

// a FilteredNoise passed where a Noise is expected samples its device table
inline DeviceNoise as_device(const FilteredNoise &n) { return n.device_noise(); }

// ---- model (src/qew_model.h) ----------------------------------------------
enum class Model { Linear = QEW_LINEAR, Arclength = QEW_ARCLENGTH };

struct Params {
    real_t physical_size = 1.0;
    real_t line_tension = 1.0;
    real_t driving_force = 0.0;
    Model model = Model::Linear;
    qew_params c() const {
        return qew_params{physical_size, line_tension, driving_force, static_cast<int>(model)};
    }
};

template <class Noise>
real_t objective(const View1D &h, const Params &p, const Noise &noise) {
    const qew_params c = p.c();
    double e = 0;
    check(qew_objective(context(), as_device(noise).handle(), &c, (int)h.extent(0), h.data(), &e));
    return e;
}

template <class Noise>
real_t objective(const View1D &x0, const View1D &d, real_t alpha, const Params &p,
                 const Noise &noise) {
    const qew_params c = p.c();
    double e = 0;
    check(qew_objective_trial(context(), as_device(noise).handle(), &c, (int)x0.extent(0),
                              x0.data(), d.data(), alpha, &e));
    return e;
}

// The per-edge quantities of gradient() (src/qew_model.h:126-147): the
// fused kernel evaluates them in registers and, when a workspace is passed,
// also writes them here, so a caller can read them after the call.
struct GradientWorkspace {
    View1D edge_dvdy;   // dV/dy at the midpoint of edge (i, i+1 mod n)
    View1D edge_slope;  // arclength only: s / sqrt(1 + s^2) on that edge

    void ensure(int n, Model model) {
        if ((int)edge_dvdy.extent(0) != n) edge_dvdy = View1D("edge_dvdy", (size_t)n);
        if (model == Model::Arclength && (int)edge_slope.extent(0) != n)
            edge_slope = View1D("edge_slope", (size_t)n);
    }
};

template <class Noise>
void gradient(const View1D &h, const Params &p, const Noise &noise, const View1D &grad) {
    if (grad.extent(0) != h.extent(0)) throw std::invalid_argument("grad extent differs from h");
    const qew_params c = p.c();
    check(qew_gradient(context(), as_device(noise).handle(), &c, (int)h.extent(0), h.data(),
                       grad.data()));
}

template <class Noise>
void gradient(const View1D &h, const Params &p, const Noise &noise, const View1D &grad,
              GradientWorkspace &ws) {
    if (grad.extent(0) != h.extent(0)) throw std::invalid_argument("grad extent differs from h");
    const int n = (int)h.extent(0);
    ws.ensure(n, p.model);
    const qew_params c = p.c();
    check(qew_gradient_edges(context(), as_device(noise).handle(), &c, n, h.data(), grad.data(),
                             ws.edge_dvdy.data(),
                             p.model == Model::Arclength ? ws.edge_slope.data() : nullptr));
}

// ---- L-BFGS (src/lbfgs.h) ---------------------------------------------------
struct LbfgsParams {
    int m = 8;
    real_t ftol = 1e-6;
    int max_iter = 10000;
    real_t c1 = 1e-4;
    real_t ls_shrink = 0.5;
    int max_ls = 30;
    real_t curvature_eps = 1e-12;
    bool precondition = false;
    real_t precond_shift = 0;
    qew_lbfgs_params c() const {
        return qew_lbfgs_params{m,      ftol,          max_iter,          c1,           ls_shrink,
                                max_ls, curvature_eps, precondition ? 1 : 0, precond_shift};
    }
};

struct LbfgsResult {
    int iterations = 0;
    real_t max_force = 0;
    bool converged = false;
    long n_func_evals = 0;
    long n_curvature_skips = 0;
    long n_ls_restarts = 0;
};

inline LbfgsResult from_c(const qew_lbfgs_result &r) {
    return LbfgsResult{r.iterations,   r.max_force,         r.converged != 0,
                       r.n_func_evals, r.n_curvature_skips, r.n_ls_restarts};
}

template <class Noise>
LbfgsResult lbfgs_minimize(const View1D &h, const Params &p, const Noise &noise,
                           const LbfgsParams &lp) {
    const qew_params c = p.c();
    const qew_lbfgs_params l = lp.c();
    qew_lbfgs_result r{};
    check(qew_lbfgs_minimize(context(), as_device(noise).handle(), &c, &l, (int)h.extent(0),
                             h.data(), &r));
    return from_c(r);
}

// ---- FIRE (src/fire.h) ----------------------------------------------------------
struct FireParams {
    real_t dt_max = 0;  // required, > 0
    real_t dt_init = 0;
    real_t ftol = 1e-6;
    int max_iter = 100000;
    int n_min = 5;
    real_t f_inc = 1.1;
    real_t f_dec = 0.5;
    real_t alpha_start = 0.1;
    real_t f_alpha = 0.99;
    qew_fire_params c() const {
        return qew_fire_params{dt_max, dt_init, ftol,        max_iter, n_min,
                               f_inc,  f_dec,   alpha_start, f_alpha};
    }
};

struct FireResult {
    int iterations = 0;
    real_t max_force = 0;
    bool converged = false;
};

template <class Noise>
FireResult fire_minimize(const View1D &h, const Params &p, const Noise &noise,
                         const FireParams &fp) {
    const qew_params c = p.c();
    const qew_fire_params f = fp.c();
    qew_fire_result r{};
    check(qew_fire_minimize(context(), as_device(noise).handle(), &c, &f, (int)h.extent(0),
                            h.data(), &r));
    return FireResult{r.iterations, r.max_force, r.converged != 0};
}

// ---- Langevin (src/dynamics.h) -------------------------------------------------
class LangevinDynamics {
public:
    LangevinDynamics(int n, std::uint64_t seed) : n_(n), seed_(seed) {}
    template <class Noise>
    void step(const View1D &h, const Params &p, const Noise &noise, real_t drift_coeff,
              real_t diff_coeff) {
        if ((int)h.extent(0) != n_) throw std::invalid_argument("line length differs");
        const qew_params c = p.c();
        check(qew_langevin_step(context(), as_device(noise).handle(), &c, n_, h.data(), drift_coeff,
                                diff_coeff, seed_, step_));
        ++step_;
    }
    // `count` steps with per-step coefficients drift[k], diff[k]; the same
    // stream and result as `count` calls of step(), several steps per launch.
    // No reference counterpart.
    template <class Noise>
    void steps(const View1D &h, const Params &p, const Noise &noise, const real_t *drift,
               const real_t *diff, int count) {
        if ((int)h.extent(0) != n_) throw std::invalid_argument("line length differs");
        if (count <= 0) return;
        const qew_params c = p.c();
        const DeviceNoise dn = as_device(noise);
        qew_noise nh = dn.handle();
        const int len = n_;
        qew_batch b = nullptr;
        check(qew_batch_create(context(), 1, &len, &nh, &c, &b));
        std::unique_ptr<qew_batch_s, qew_status (*)(qew_batch)> owner(b, qew_batch_destroy);
        check(qew_batch_langevin(b, QEW_F64, h.data(), nullptr, count, drift, diff, seed_, step_));
        step_ += (std::uint64_t)count;
    }
    std::uint64_t steps_taken() const { return step_; }

private:
    int n_;
    std::uint64_t seed_;
    std::uint64_t step_ = 0;
};

// ---- Rosso-Krauth (src/rosso_krauth.h) -------------------------------------------
struct RkParams {
    real_t dstep = 0;
    real_t max_advance = 0;
    real_t root_tol = 1e-10;
    real_t sweep_tol = 1e-9;
    real_t runaway = 0;
    int max_sweeps = 200000;
    qew_rk_params c() const {
        return qew_rk_params{dstep, max_advance, root_tol, sweep_tol, runaway, max_sweeps};
    }
};

struct RkResult {
    bool blocked = false;
    bool runaway = false;
    int sweeps = 0;
    real_t total_advance = 0;
    real_t mean_h = 0;
    // site-velocity evaluations the reference algorithm makes (no reference field)
    long n_velocity_evals = 0;
    // evaluations performed on the device (unchanged sites skipped)
    long n_velocity_evals_done = 0;
};

inline RkResult from_c(const qew_rk_result &r) {
    return RkResult{r.blocked != 0, r.runaway != 0, r.sweeps, r.total_advance, r.mean_h,
                    r.n_velocity_evals, r.n_velocity_evals_done};
}

namespace detail {
// qew_params whose physical_size / n reproduces dx exactly (the kernels
// derive the spacing as L / n).
inline qew_params params_for_spacing(int n, real_t dx, real_t lt, real_t f, Model model) {
    double L = dx * (double)n;
    for (int k = 0; k < 8 && L / (double)n != dx; ++k)
        L = std::nextafter(L, (L / (double)n < dx) ? HUGE_VAL : -HUGE_VAL);
    if (L / (double)n != dx) throw std::invalid_argument("no box length L with L / n == dx");
    return qew_params{L, lt, f, static_cast<int>(model)};
}
}  // namespace detail

// rk_site_velocity (src/rosso_krauth.h:52-82) at sites i[k] of a line of n
// sites of spacing dx, trial heights hi[k], neighbours hl[k], hr[k] fixed;
// v[k] out. One launch for all sites.
template <class Noise>
void rk_site_velocity(const std::vector<int> &i, int n, real_t dx, real_t lt, real_t f,
                      Model model, const View1D &hi, const View1D &hl, const View1D &hr,
                      const Noise &noise, const View1D &v) {
    const size_t m = i.size();
    if (hi.extent(0) != m || hl.extent(0) != m || hr.extent(0) != m || v.extent(0) != m)
        throw std::invalid_argument("rk_site_velocity: views of different extents");
    const qew_params c = detail::params_for_spacing(n, dx, lt, f, model);
    const DeviceInts idx(i);
    check(qew_rk_site_velocity(context(), as_device(noise).handle(), &c, n, (int)m, idx.data(),
                               hi.data(), hl.data(), hr.data(), v.data()));
}

// The same for one site, with the reference's argument list (one launch).
template <class Noise>
real_t rk_site_velocity(int i, int n, real_t dx, real_t lt, real_t f, Model model, real_t hi,
                        real_t hl, real_t hr, const Noise &noise) {
    View1D H(1), Hl(1), Hr(1), V(1);
    H.copy_from_host(&hi);
    Hl.copy_from_host(&hl);
    Hr.copy_from_host(&hr);
    rk_site_velocity(std::vector<int>{i}, n, dx, lt, f, model, H, Hl, Hr, noise, V);
    return V.to_host()[0];
}

// rk_forward_zero (src/rosso_krauth.h:88-124) at sites i[k]: out[k] the
// nearest forward zero of the site's velocity (hi[k] itself when it is
// <= 0), found[k] false when the march reached rp.max_advance without one.
template <class Noise>
void rk_forward_zero(const std::vector<int> &i, int n, real_t dx, real_t lt, real_t f,
                     Model model, const View1D &hi, const View1D &hl, const View1D &hr,
                     const Noise &noise, const RkParams &rp, const View1D &out,
                     std::vector<bool> &found) {
    const size_t m = i.size();
    if (hi.extent(0) != m || hl.extent(0) != m || hr.extent(0) != m || out.extent(0) != m)
        throw std::invalid_argument("rk_forward_zero: views of different extents");
    const qew_params c = detail::params_for_spacing(n, dx, lt, f, model);
    const qew_rk_params r = rp.c();
    const DeviceInts idx(i);
    const DeviceInts flags(std::vector<int>(m, 0));
    check(qew_rk_forward_zero(context(), as_device(noise).handle(), &c, &r, n, (int)m, idx.data(),
                              hi.data(), hl.data(), hr.data(), out.data(),
                              const_cast<int *>(flags.data()), nullptr));
    std::vector<int> fh(m);
    if (m) check(qew_copy_to_host(context(), fh.data(), flags.data(), sizeof(int) * m));
    found.assign(m, false);
    for (size_t k = 0; k < m; ++k) found[k] = fh[k] != 0;
}

// The same for one site, with the reference's argument list.
template <class Noise>
real_t rk_forward_zero(int i, int n, real_t dx, real_t lt, real_t f, Model model, real_t hi,
                       real_t hl, real_t hr, const Noise &noise, const RkParams &rp,
                       bool &found) {
    View1D H(1), Hl(1), Hr(1), O(1);
    H.copy_from_host(&hi);
    Hl.copy_from_host(&hl);
    Hr.copy_from_host(&hr);
    std::vector<bool> fd;
    rk_forward_zero(std::vector<int>{i}, n, dx, lt, f, model, H, Hl, Hr, noise, rp, O, fd);
    found = fd[0];
    return O.to_host()[0];
}

template <class Noise>
RkResult rk_block(const View1D &h, const Params &p, const Noise &noise, const RkParams &rp) {
    const qew_params c = p.c();
    const qew_rk_params r = rp.c();
    qew_rk_result out{};
    check(qew_rk_block(context(), as_device(noise).handle(), &c, &r, (int)h.extent(0), h.data(),
                       &out));
    return from_c(out);
}

}  // namespace qew

#endif  // QEW_FACADE_HPP
