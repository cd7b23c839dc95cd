// lbfgs_minimize for one line in FP64, iterations driven from the host with
// the reference's control flow (src/lbfgs.h:94-287): every dot product and
// max-norm returns to the host, every vector operation is its own launch. This
// is the plain single-line GPU path; the batched persistent minimiser lives in
// qew_lbfgs_batch.hip. Vector kernels: reductions finish in the last
// workgroup (fixed order, deterministic); the FFT preconditioner is
// d = irfft(rfft(q) / eig) computed as a complex transform of the real vector
// with the Hermitian-symmetric divisor eig[min(k, n - k)].
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kRedBlocks = 512;

template <typename T>
__device__ __forceinline__ T ldv(const T *p) {
    return *reinterpret_cast<const volatile T *>(p);
}

// NV values per element, reduced into out[0..NV) (device). op 0: sums of
// products given by the functor, op 1: max.
template <int NV, class F>
__device__ void reduce_to(int n, F fn, double *part, unsigned *counter, double *out, bool is_max) {
    __shared__ double red[16];
    __shared__ int s_last;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double x[NV];
        fn(i, x);
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = is_max ? (x[v] > acc[v] ? x[v] : acc[v]) : acc[v] + x[v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double r = is_max ? block_max(acc[v], red) : block_sum(acc[v], red);
        if (threadIdx.x == 0) part[v * gridDim.x + blockIdx.x] = r;
    }
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double a = 0;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
            const double p = ldv(&part[v * gridDim.x + b]);
            a = is_max ? (p > a ? p : a) : a + p;
        }
        a = is_max ? block_max(a, red) : block_sum(a, red);
        if (threadIdx.x == 0) out[v] = a;
    }
    if (threadIdx.x == 0) *counter = 0u;
}

__global__ __launch_bounds__(256) void k_dot(int n, const double *a, const double *b, double *part,
                                             unsigned *counter, double *out) {
    reduce_to<1>(n, [&](int i, double *x) { x[0] = a[i] * b[i]; }, part, counter, out, false);
}

__global__ __launch_bounds__(256) void k_maxabs(int n, const double *a, double *part,
                                                unsigned *counter, double *out) {
    reduce_to<1>(n, [&](int i, double *x) { x[0] = fabs(a[i]); }, part, counter, out, true);
}

// s = al d, y = g_new - g written out; sy, ss, yy reduced (src/lbfgs.h:250-261)
__global__ __launch_bounds__(256) void k_sy(int n, double al, const double *d, const double *gn,
                                            const double *g, double *s, double *y, double *part,
                                            unsigned *counter, double *out) {
    reduce_to<3>(n, [&](int i, double *x) {
        const double si = al * d[i];
        const double yi = gn[i] - g[i];
        s[i] = si;
        y[i] = yi;
        x[0] = si * yi;
        x[1] = si * si;
        x[2] = yi * yi;
    }, part, counter, out, false);
}

__global__ void k_axpy(int n, double c, const double *x, double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += c * x[i];
}

__global__ void k_scale(int n, double c, const double *x, double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = c * x[i];
}

__global__ void k_neg(int n, double *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = -d[i];
}

__global__ void k_pack(int n, const double *q, double2 *z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = make_double2(q[i], 0.0);
}

// eig[k] for k = 0..n/2 (src/lbfgs.h:138-143)
__global__ void k_eig(int nh, int n, double kap, double dx, double sh, double *eig) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh) return;
    const double pi = 3.14159265358979323846;
    const double s = sin(pi * (double)k / (double)n);
    eig[k] = (kap / dx) * 4 * s * s + sh;
}

__global__ void k_divide_eig(int n, double2 *z, const double *eig) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double e = eig[k <= n / 2 ? k : n - k];
    z[k] = make_double2(z[k].x / e, z[k].y / e);
}

__global__ void k_unpack(int n, const double2 *z, double *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = z[i].x / (double)n;
}

dim3 g1(int n) { return dim3((n + 255) / 256); }

struct Plain {
    qew_ctx ctx;
    hipStream_t s;
    int n;
    double *part;
    unsigned *counter;
    double *out;   // device, 4 doubles
    double *host;  // pinned, 4 doubles

    int red_blocks() const { return std::max(1, std::min(kRedBlocks, (n + 255) / 256)); }

    double fetch(int count) {
        QEW_HIP(hipMemcpyAsync(host, out, sizeof(double) * count, hipMemcpyDeviceToHost, s));
        QEW_HIP(hipStreamSynchronize(s));
        return host[0];
    }
    double dot(const double *a, const double *b) {
        hipLaunchKernelGGL(k_dot, dim3(red_blocks()), dim3(256), 0, s, n, a, b, part, counter, out);
        QEW_HIP(hipGetLastError());
        return fetch(1);
    }
    double max_abs(const double *a) {
        hipLaunchKernelGGL(k_maxabs, dim3(red_blocks()), dim3(256), 0, s, n, a, part, counter, out);
        QEW_HIP(hipGetLastError());
        return fetch(1);
    }
    void axpy(double c, const double *x, double *y) {
        hipLaunchKernelGGL(k_axpy, g1(n), dim3(256), 0, s, n, c, x, y);
    }
};

}  // namespace

using namespace qew_host;

extern "C" {

qew_lbfgs_params qew_lbfgs_default_params(void) {
    qew_lbfgs_params p;
    p.m = 8;
    p.ftol = 1e-6;
    p.max_iter = 10000;
    p.c1 = 1e-4;
    p.ls_shrink = 0.5;
    p.max_ls = 30;
    p.curvature_eps = 1e-12;
    p.precondition = 0;
    p.precond_shift = 0;
    return p;
}

qew_status qew_lbfgs_minimize(qew_ctx ctx, qew_noise noise, const qew_params *p,
                              const qew_lbfgs_params *lp, int n, double *h,
                              qew_lbfgs_result *result) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && lp && h && result, "null pointer");
        QEW_REQUIRE(n >= 1, "line length must be >= 1");
        QEW_REQUIRE(lp->m >= 1, "history length m must be >= 1");
        require_wide_table(noise->desc);
        use_device(ctx);
        const int m = lp->m;
        const LineDesc line = single_line(n, p);
        ctx_reserve_eval(ctx, line.ntiles, 1);
        // workspace: g, g_new, d, q, s_tmp, y_tmp, S[m], Y[m], eig, 2 complex
        // FFT buffers, reduction partials and the energy slot
        const size_t nn = (size_t)n;
        const size_t nvec = 6 + 2 * (size_t)m;
        const size_t bytes = sizeof(double) * (nvec * nn + (nn / 2 + 2) + 4 * nn + 3 * kRedBlocks + 16) +
                             256;
        ctx->work.ensure(bytes);
        double *base = ctx->work.as<double>();
        double *g = base, *g_new = g + nn, *d = g_new + nn, *q = d + nn, *s_tmp = q + nn,
               *y_tmp = s_tmp + nn;
        std::vector<double *> S(m), Y(m);
        for (int k = 0; k < m; ++k) {
            S[k] = y_tmp + nn * (1 + 2 * k);
            Y[k] = S[k] + nn;
        }
        double *eig = y_tmp + nn * (1 + 2 * m);
        double2 *z = reinterpret_cast<double2 *>(eig + ((nn / 2 + 2) / 2) * 2);
        double2 *zt = z + nn;
        double *part = reinterpret_cast<double *>(zt + nn);
        double *out = part + 3 * kRedBlocks;
        unsigned *counter = reinterpret_cast<unsigned *>(out + 8);
        hipStream_t s = ctx->stream;
        QEW_HIP(hipMemsetAsync(counter, 0, 64, s));
        Plain P{ctx, s, n, part, counter, out, ctx->pinned.as<double>()};

        auto eval = [&](const double *x, const double *dir, double alpha, double *grad,
                        bool energy) -> double {
            EvalArgs a{};
            a.ntiles = line.ntiles;
            a.h = x;
            a.d = dir;
            a.alpha = alpha;
            a.grad = grad;
            a.energy = energy ? ctx->result.as<double>() : nullptr;
            a.partial = ctx->partial.ptr;
            launch_eval_f64(s, a, &line, &noise->desc);
            if (!energy) return 0.0;
            QEW_HIP(hipMemcpyAsync(P.host, ctx->result.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
            QEW_HIP(hipStreamSynchronize(s));
            return P.host[0];
        };

        const int nh = n / 2 + 1;
        if (lp->precondition) {
            const double kappa = p->line_tension;
            const double dx = p->physical_size / (double)n;
            const double pi = 3.14159265358979323846;
            double shift = lp->precond_shift;
            if (shift <= 0) {
                const double s1 = std::sin(pi / (double)n);
                shift = (kappa / dx) * 4 * s1 * s1;
            }
            hipLaunchKernelGGL(k_eig, g1(nh), dim3(256), 0, s, nh, n, kappa, dx, shift, eig);
            QEW_HIP(hipGetLastError());
        }

        std::vector<double> rho(m, 0.0);
        std::vector<int> slots;
        double gamma = 1.0;
        eval(h, nullptr, 0.0, g, false);
        qew_lbfgs_result res{};
        int ls_fail_streak = 0;
        double phi = 0;
        bool done = false;
        for (int it = 0; it < lp->max_iter && !done; ++it) {
            const double gmax = P.max_abs(g);
            if (gmax < lp->ftol) {
                res.iterations = it;
                res.max_force = gmax;
                res.converged = 1;
                done = true;
                break;
            }
            QEW_HIP(hipMemcpyAsync(q, g, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
            const int hs = (int)slots.size();
            std::vector<double> la(hs);
            for (int idx = hs - 1; idx >= 0; --idx) {
                const int row = slots[idx];
                const double a = rho[row] * P.dot(S[row], q);
                la[idx] = a;
                P.axpy(-a, Y[row], q);
            }
            if (lp->precondition) {
                hipLaunchKernelGGL(k_pack, g1(n), dim3(256), 0, s, n, q, z);
                fft_c2c(s, z, zt, n, 1, 1, n, -1);
                hipLaunchKernelGGL(k_divide_eig, g1(n), dim3(256), 0, s, n, z, eig);
                fft_c2c(s, z, zt, n, 1, 1, n, +1);
                hipLaunchKernelGGL(k_unpack, g1(n), dim3(256), 0, s, n, z, d);
            } else {
                const double g0 = slots.empty() ? 1.0 : gamma;
                hipLaunchKernelGGL(k_scale, g1(n), dim3(256), 0, s, n, g0, q, d);
            }
            for (int idx = 0; idx < hs; ++idx) {
                const int row = slots[idx];
                const double b = rho[row] * P.dot(Y[row], d);
                P.axpy(la[idx] - b, S[row], d);
            }
            hipLaunchKernelGGL(k_neg, g1(n), dim3(256), 0, s, n, d);
            double dphi0 = P.dot(g, d);
            if (dphi0 >= 0) {
                slots.clear();
                hipLaunchKernelGGL(k_scale, g1(n), dim3(256), 0, s, n, -1.0, g, d);
                dphi0 = -P.dot(g, g);
            }
            if (it == 0) {
                phi = eval(h, nullptr, 0.0, nullptr, true);
                ++res.n_func_evals;
            }
            const double phi0 = phi;
            double alpha = 1.0;
            if (!lp->precondition && slots.empty())
                alpha = 1.0 / std::max(1.0, std::sqrt(P.dot(g, g)));
            bool ok = false;
            for (int ls = 0; ls < lp->max_ls; ++ls) {
                const double phi_trial = eval(h, d, alpha, nullptr, true);
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
                    res.iterations = it;
                    res.max_force = gmax;
                    res.converged = 0;
                    done = true;
                    break;
                }
                slots.clear();
                ++res.n_ls_restarts;
                continue;
            }
            ls_fail_streak = 0;
            const double al = alpha;
            P.axpy(al, d, h);
            eval(h, nullptr, 0.0, g_new, false);
            hipLaunchKernelGGL(k_sy, dim3(P.red_blocks()), dim3(256), 0, s, n, al, d, g_new, g, s_tmp,
                               y_tmp, part, counter, out);
            QEW_HIP(hipGetLastError());
            P.fetch(3);
            const double sy = P.host[0], ss = P.host[1], yy = P.host[2];
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
        if (!done) {
            res.iterations = lp->max_iter;
            res.max_force = P.max_abs(g);
            res.converged = 0;
        }
        QEW_HIP(hipStreamSynchronize(s));
        *result = res;
    });
}

}  // extern "C"
