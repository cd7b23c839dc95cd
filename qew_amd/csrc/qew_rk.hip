// Rosso-Krauth no-passing construction (src/rosso_krauth.h:52-194) for
// batches of lines. Each colour pass is one launch over every still-active
// line (one thread per site of that colour: its neighbours are the other
// colour, so the sites are independent); a per-line reduction then updates
// the line's sweep status on the device (runaway / blocked / running), so
// finished lines freeze while the others continue and the host reads back
// only an "any line still running" flag every few sweeps. The line state,
// the velocity sum, the v <= 0 tests and the bisection are FP64; the tables
// may be stored as FP64 or FP32.
#include <algorithm>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kRkBlock = 256;

struct RkState {
    int active;
    int sweeps;
    int blocked;
    int runaway;
    double cum_adv;
    double total_adv;
    double mean_h0;
    double mean_h;
};

struct RkArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    double *h;
    double *adv;
    RkState *state;
    qew_rk_params rp;
};

// rk_site_velocity (src/rosso_krauth.h:52-82)
template <typename C>
__device__ __forceinline__ double site_velocity(int i, int n, double dx, double lt, double f,
                                                int model, double hi, double hl, double hr,
                                                const NoiseDesc &nd) {
    const int im = (i == 0) ? n - 1 : i - 1;
    const int ip = (i + 1 == n) ? 0 : i + 1;
    double line_force;
    if (model == 0) {
        line_force = -lt * (hl + hr - 2.0 * hi) / (dx * dx);
    } else {
        const double dhr = (hr - hi) / dx;
        const double dhl = (hi - hl) / dx;
        line_force = lt * (dhl / sqrt(1.0 + dhl * dhl) - dhr / sqrt(1.0 + dhr * dhr)) / dx;
    }
    const double xcl = (double)(im + i) * 0.5 * dx;
    const double hcl = (hl + hi) * 0.5;
    const double xcr = (double)(i + ip) * 0.5 * dx;
    const double hcr = (hi + hr) * 0.5;
    const double dvl = sample_dy64<C>(nd, xcl, hcl);
    const double dvr = sample_dy64<C>(nd, xcr, hcr);
    const double grad = dx * (line_force + 0.5 * (dvl + dvr) - f);
    return -grad;
}

// rk_forward_zero (src/rosso_krauth.h:88-124): march in dstep steps
// (accumulating a = b), then at most 60 bisections.
template <typename C>
__device__ double forward_zero(int i, int n, double dx, double lt, double f, int model, double hi,
                               double hl, double hr, const NoiseDesc &nd, const qew_rk_params &rp) {
    const double v0 = site_velocity<C>(i, n, dx, lt, f, model, hi, hl, hr, nd);
    if (v0 <= 0) return hi;
    double a = hi;
    double advanced = 0;
    while (advanced < rp.max_advance) {
        const double b = a + rp.dstep;
        const double vb = site_velocity<C>(i, n, dx, lt, f, model, b, hl, hr, nd);
        if (vb <= 0) {
            double lo = a, hic = b;
            for (int k = 0; k < 60 && (hic - lo) > rp.root_tol; ++k) {
                const double mid = 0.5 * (lo + hic);
                const double vm = site_velocity<C>(i, n, dx, lt, f, model, mid, hl, hr, nd);
                if (vm > 0) lo = mid;
                else hic = mid;
            }
            return 0.5 * (lo + hic);
        }
        a = b;
        advanced += rp.dstep;
    }
    return a;
}

template <typename C>
__global__ __launch_bounds__(kRkBlock) void rk_color_kernel(RkArgs a, int color) {
    const int line = blockIdx.y;
    if (!a.state[line].active) return;
    const LineDesc L = a.lines[line];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 2 * j + color;
    if (i >= L.n) return;
    const NoiseDesc nd = a.noises[L.noise];
    double *h = a.h + L.off;
    const int n = L.n;
    const int im = (i == 0) ? n - 1 : i - 1;
    const int ip = (i + 1 == n) ? 0 : i + 1;
    const double dx = L.L / (double)n;
    const double hi = h[i];
    const double hnew = forward_zero<C>(i, n, dx, L.lt, L.f, L.model, hi, h[im], h[ip], nd, a.rp);
    a.adv[L.off + i] = hnew - hi;
    h[i] = hnew;
}

// One workgroup per line: sum of the sweep's advances, then the
// runaway/blocked decision (src/rosso_krauth.h:168-188).
__global__ __launch_bounds__(kRkBlock) void rk_status_kernel(RkArgs a, unsigned *running) {
    __shared__ double red[16];
    const int line = blockIdx.x;
    RkState st = a.state[line];
    if (!st.active) return;
    const LineDesc L = a.lines[line];
    double acc = 0;
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) acc += a.adv[L.off + i];
    const double total = block_sum(acc, red);
    if (threadIdx.x != 0) return;
    st.cum_adv += total;
    const double mean_adv = st.cum_adv / (double)L.n;
    st.sweeps += 1;
    st.total_adv = total;
    st.mean_h = st.mean_h0 + mean_adv;
    if (mean_adv > a.rp.runaway) {
        st.runaway = 1;
        st.blocked = 0;
        st.active = 0;
    } else if (total < a.rp.sweep_tol) {
        st.blocked = 1;
        st.runaway = 0;
        st.active = 0;
    }
    if (st.active && st.sweeps >= a.rp.max_sweeps) st.active = 0;  // timeout
    a.state[line] = st;
    if (st.active) atomicAdd(running, 1u);
}

__global__ __launch_bounds__(kRkBlock) void rk_init_kernel(RkArgs a) {
    __shared__ double red[16];
    const int line = blockIdx.x;
    const LineDesc L = a.lines[line];
    double acc = 0;
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) acc += a.h[L.off + i];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
        RkState st{};
        st.active = a.rp.max_sweeps > 0 ? 1 : 0;
        st.mean_h0 = s / (double)L.n;
        st.mean_h = 0;
        a.state[line] = st;
    }
}

__global__ void rk_velocity_kernel(NoiseDesc nd, int n, double dx, double lt, double f, int model,
                                   int m, const int *idx, const double *hi, const double *hl,
                                   const double *hr, double *v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    if (nd.dtype == 1)
        v[k] = site_velocity<float>(idx[k], n, dx, lt, f, model, hi[k], hl[k], hr[k], nd);
    else
        v[k] = site_velocity<double>(idx[k], n, dx, lt, f, model, hi[k], hl[k], hr[k], nd);
}

// Runs the sweeps for `nlines` lines whose descriptors are on the device.
void run_rk(qew_ctx ctx, const LineDesc *d_lines, const NoiseDesc *d_noises, int nlines, int max_n,
            int table_dtype, const qew_rk_params &rp, double *h, double *adv, RkState *state,
            unsigned *running, unsigned *running_host, std::vector<RkState> &out) {
    // a march step of zero (or NaN) with a positive advance cap never ends;
    // the reference would spin forever, the device must not
    QEW_REQUIRE(rp.dstep > 0 || !(rp.max_advance > 0),
                "rk_block: dstep must be positive when max_advance > 0");
    hipStream_t s = ctx->stream;
    RkArgs a{d_lines, d_noises, h, adv, state, rp};
    hipLaunchKernelGGL(rk_init_kernel, dim3(nlines), dim3(kRkBlock), 0, s, a);
    QEW_HIP(hipGetLastError());
    const int per_color = (max_n + 1) / 2;
    const dim3 grid((per_color + kRkBlock - 1) / kRkBlock, nlines);
    // check the running flag every `poll` sweeps; lines that finished in
    // between are frozen, so the extra launches are no-ops for them
    int poll = 4;
    for (int sweep = 0; sweep < rp.max_sweeps;) {
        QEW_HIP(hipMemsetAsync(running, 0, sizeof(unsigned), s));
        const int todo = std::min(poll, rp.max_sweeps - sweep);
        for (int k = 0; k < todo; ++k) {
            for (int color = 0; color < 2; ++color) {
                if (table_dtype == 1)
                    hipLaunchKernelGGL(rk_color_kernel<float>, grid, dim3(kRkBlock), 0, s, a, color);
                else
                    hipLaunchKernelGGL(rk_color_kernel<double>, grid, dim3(kRkBlock), 0, s, a, color);
            }
            if (k == todo - 1) QEW_HIP(hipMemsetAsync(running, 0, sizeof(unsigned), s));
            hipLaunchKernelGGL(rk_status_kernel, dim3(nlines), dim3(kRkBlock), 0, s, a, running);
        }
        QEW_HIP(hipGetLastError());
        sweep += todo;
        QEW_HIP(hipMemcpyAsync(running_host, running, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        QEW_HIP(hipStreamSynchronize(s));
        if (*running_host == 0) break;
        poll = std::min(poll * 2, 256);
    }
    out.resize(nlines);
    QEW_HIP(hipMemcpyAsync(out.data(), state, sizeof(RkState) * nlines, hipMemcpyDeviceToHost, s));
    QEW_HIP(hipStreamSynchronize(s));
}

qew_rk_result to_result(const RkState &st) {
    qew_rk_result r{};
    r.blocked = st.blocked;
    r.runaway = st.runaway;
    r.sweeps = st.sweeps;
    r.total_advance = st.total_adv;
    r.mean_h = st.sweeps > 0 ? st.mean_h : 0.0;
    return r;
}

}  // namespace

using namespace qew_host;

extern "C" {

qew_status qew_rk_block(qew_ctx ctx, qew_noise noise, const qew_params *p,
                        const qew_rk_params *rp, int n, double *h, qew_rk_result *result) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && rp && h && result, "null pointer");
        QEW_REQUIRE(n >= 1, "line length must be >= 1");
        if (n % 2 != 0)
            throw Error(QEW_ERR_INVALID_ARGUMENT,
                        "rk_block: line length must be even for the red-black sweep");
        use_device(ctx);
        LineDesc l = single_line(n, p);
        // workspace: [LineDesc][NoiseDesc][RkState][running][adv n]
        const size_t off_noise = 256, off_state = 512, off_run = 768, off_adv = 1024;
        ctx->work.ensure(off_adv + sizeof(double) * (size_t)n);
        char *w = ctx->work.as<char>();
        ctx->pinned.ensure(sizeof(double) * 64);
        QEW_HIP(hipMemcpyAsync(w, &l, sizeof(LineDesc), hipMemcpyHostToDevice, ctx->stream));
        QEW_HIP(hipMemcpyAsync(w + off_noise, &noise->desc, sizeof(NoiseDesc), hipMemcpyHostToDevice,
                               ctx->stream));
        std::vector<RkState> out;
        const int dt = noise->desc.kind == TABLE ? noise->desc.dtype : 0;
        run_rk(ctx, reinterpret_cast<LineDesc *>(w), reinterpret_cast<NoiseDesc *>(w + off_noise), 1, n,
               dt, *rp, h, reinterpret_cast<double *>(w + off_adv),
               reinterpret_cast<RkState *>(w + off_state), reinterpret_cast<unsigned *>(w + off_run),
               ctx->pinned.as<unsigned>(), out);
        *result = to_result(out[0]);
    });
}

qew_status qew_batch_rk_block(qew_batch b, const qew_rk_params *rp, double *h,
                              qew_rk_result *results) {
    return guard([&] {
        QEW_REQUIRE(b && rp && h && results, "null pointer");
        for (const auto &l : b->lines)
            if (l.n % 2 != 0)
                throw Error(QEW_ERR_INVALID_ARGUMENT,
                            "rk_block: line length must be even for the red-black sweep");
        for (const auto &nd : b->noises)
            QEW_REQUIRE(nd.kind != TABLE || nd.dtype == b->noises[0].dtype,
                        "all tables of a batch must share one storage type");
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        const size_t off_state = 0;
        const size_t off_run = ((sizeof(RkState) * b->nlines + 255) / 256) * 256;
        const size_t off_adv = off_run + 256;
        b->d_scratch.ensure(off_adv + sizeof(double) * (size_t)b->total);
        ctx->pinned.ensure(sizeof(double) * 64);
        char *w = b->d_scratch.as<char>();
        std::vector<RkState> out;
        const int dt = b->noises[0].kind == TABLE ? b->noises[0].dtype : 0;
        run_rk(ctx, b->d_lines.as<LineDesc>(), b->d_noises.as<NoiseDesc>(), b->nlines, b->max_n, dt,
               *rp, h, reinterpret_cast<double *>(w + off_adv),
               reinterpret_cast<RkState *>(w + off_state), reinterpret_cast<unsigned *>(w + off_run),
               ctx->pinned.as<unsigned>(), out);
        for (int l = 0; l < b->nlines; ++l) results[l] = to_result(out[l]);
    });
}

qew_status qew_rk_site_velocity(qew_ctx ctx, qew_noise noise, const qew_params *p, int n, int m,
                                const int *i, const double *hi, const double *hl,
                                const double *hr, double *v) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p, "null pointer");
        QEW_REQUIRE(n >= 1 && m >= 0, "bad sizes");
        if (m == 0) return;
        QEW_REQUIRE(i && hi && hl && hr && v, "null pointer");
        use_device(ctx);
        hipLaunchKernelGGL(rk_velocity_kernel, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                           noise->desc, n, p->physical_size / (double)n, p->line_tension,
                           p->driving_force, p->model, m, i, hi, hl, hr, v);
        QEW_HIP(hipGetLastError());
    });
}

}  // extern "C"
