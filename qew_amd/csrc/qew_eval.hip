// Fused gradient + energy of ragged batches of lines (the qew_model.h hot
// path: src/qew_model.h:53-201). One workgroup of 256 threads = one tile of
// 255 consecutive sites of one line: thread t >= 1 owns site first - 1 + t
// and its right edge, thread 0 evaluates the edge left of the first site (the
// halo), so every wave makes one round of 4x4 gathers. The left-neighbour
// edge of each site comes through LDS. Each tile writes
// its energy partial; a second, small launch sums every line's partials in a
// fixed order (deterministic). The kernel boundary orders the two: no
// device-scope fence or atomic is needed per tile, which matters on a
// multi-XCD part where a device-scope release writes back the L2.
#include "qew_internal.hpp"

#include <mutex>

using namespace qew_dev;

// Scratch of qew_eval_noalloc: the per-tile energy partials of one line, in a
// file-scope device array instead of a context buffer, so that the entry point
// allocates nothing. Every device the library runs on holds its own copy until
// the process ends. On the host this name is not a device address: host code
// uses only the per-device address hipGetSymbolAddress returns for it.
constexpr int kNoAllocMaxTiles = 16384;  // lines of up to 16384 * 255 sites
static __device__ __align__(256) double scratch[kNoAllocMaxTiles];

namespace {

// energy[line] = sum of the line's tile partials (FP64)
__global__ __launch_bounds__(256) void line_energy64(const LineDesc *lines, LineDesc one,
                                                     int single, const double *partial,
                                                     double *energy) {
    __shared__ double red[16];
    const int line = blockIdx.x;
    const LineDesc L = single ? one : lines[line];
    const int t0 = single ? 0 : L.tile0;
    double acc = 0;
    for (int t = threadIdx.x; t < L.ntiles; t += blockDim.x) acc += partial[t0 + t];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) energy[line] = acc;
}

// FP32 path: compensated partials; -f * anchor * L adds the anchor's share of
// the driving-force term
__global__ __launch_bounds__(256) void line_energy32(const LineDesc *lines, const double *anchor,
                                                     const KSum *partial, double *energy) {
    __shared__ KSum red[16];
    const int line = blockIdx.x;
    const LineDesc L = lines[line];
    KSum acc{0.f, 0.f};
    for (int t = threadIdx.x; t < L.ntiles; t += blockDim.x) acc = ksum_add(acc, partial[L.tile0 + t]);
    acc = block_ksum(acc, red);
    if (threadIdx.x == 0) {
        const double a = anchor ? anchor[line] : 0.0;
        energy[line] = ((double)acc.s + (double)acc.c) - L.f * a * L.L;
    }
}

template <typename C, bool TRIAL, bool GRAD, bool ENERGY>
__global__ __launch_bounds__(kTile) void eval64_kernel(qew_host::EvalArgs a, LineDesc one_line,
                                                       NoiseDesc one_noise, int single) {
    __shared__ double s_dvdy[kTile + 1];
    __shared__ double s_slope[kTile + 1];
    __shared__ double s_h[kTile + 1];
    __shared__ double red[16];
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    int line;
    LineDesc L;
    NoiseDesc N;
    if (single) {
        line = 0;
        L = one_line;
        N = one_noise;
        L.tile0 = 0;
    } else {
        line = a.tile_line[tile];
        L = a.lines[line];
        N = a.noises[L.noise];
    }
    const int n = L.n;
    const int first = (tile - L.tile0) * kTileSites;
    const double *h = static_cast<const double *>(a.h) + L.off;
    const double *dd = TRIAL ? static_cast<const double *>(a.d) + L.off : nullptr;
    const double alpha = a.alpha;
    auto at = [&](int i) -> double { return TRIAL ? h[i] + alpha * dd[i] : h[i]; };
    const double dx = L.L / (double)n;
    const double lt = L.lt, f = L.f;
    const bool arc = (L.model == 1);
    // this thread's edge is (i, i + 1 mod n), i = first - 1 + tid; it owns
    // site i for tid >= 1, and thread 0's edge is the halo the gradient needs
    const int gl = first - 1 + tid;
    const bool owns = tid >= 1 && gl < n;
    const bool has_edge = owns || (GRAD && tid == 0);
    const int i = gl < 0 ? n - 1 : gl;
    double e = 0, hi = 0, hip = 0;
    if (has_edge) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        hi = at(i);
        hip = at(ip);
        const double xc = (double)(i + ip) * 0.5 * dx;
        const double hc = (hip + hi) * 0.5;
        double v = 0, dvdy = 0;
        if (ENERGY && GRAD) sample_v_dy64<C>(N, xc, hc, v, dvdy);
        else if (ENERGY) v = sample_v64<C>(N, xc, hc);
        else dvdy = sample_dy64<C>(N, xc, hc);
        const double dh = (hip - hi) / dx;
        if (ENERGY && owns) {
            double le;
            if (!arc) le = lt * dh * dh * 0.5;
            else le = lt * (sqrt(1.0 + dh * dh) - 1.0);
            e = dx * (le + v - f * hi);
        }
        if (GRAD) {
            s_dvdy[tid] = dvdy;
            s_h[tid] = hi;
            if (arc) s_slope[tid] = dh / sqrt(1.0 + dh * dh);
            // the per-edge quantities of the reference's GradientWorkspace,
            // edge i written by the thread that owns site i
            if (owns && a.edge_dvdy) {
                a.edge_dvdy[L.off + i] = dvdy;
                if (arc && a.edge_slope) a.edge_slope[L.off + i] = s_slope[tid];
            }
        }
    }
    if (GRAD) {
        __syncthreads();
        if (owns) {
            double line_force;
            if (!arc) {
                const double d2h = s_h[tid - 1] + hip - 2.0 * hi;
                line_force = -lt * d2h / (dx * dx);
            } else {
                line_force = lt * (s_slope[tid - 1] - s_slope[tid]) / dx;
            }
            static_cast<double *>(a.grad)[L.off + i] =
                dx * (line_force + 0.5 * (s_dvdy[tid - 1] + s_dvdy[tid]) - f);
        }
    }
    if (ENERGY) {
        const double bsum = block_sum(e, red);
        if (tid == 0) static_cast<double *>(a.partial)[tile] = bsum;
    }
}

// FP32 fast path: FP32 line offsets from a per-line anchor, FP32 table, FP32
// arithmetic, energy accumulated as compensated (sum, error) pairs.
__global__ __launch_bounds__(kTile) void eval32_kernel(qew_host::EvalArgs a) {
    __shared__ float s_dvdy[kTile + 1];
    __shared__ float s_slope[kTile + 1];
    __shared__ float s_h[kTile + 1];
    __shared__ KSum red[16];
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int line = a.tile_line[tile];
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const int first = (tile - L.tile0) * kTileSites;
    const float *h = static_cast<const float *>(a.h) + L.off;
    const double anchor = a.anchor ? a.anchor[line] : 0.0;
    const LineGeom G = make_geom(L, N, anchor);
    const float dxf = G.dxf, inv_dx = G.inv_dx;
    const float lt = (float)L.lt, f = (float)L.f;
    const bool arc = (L.model == 1);
    auto edge = [&](int e, int ep, float he, float hep, float &v, float &dvdy) {
        edge32<true, true>(G, e, ep, he, hep, v, dvdy);
    };

    // edge (i, i + 1 mod n), i = first - 1 + tid; site i is this thread's
    // for tid >= 1, thread 0's edge is the tile's left halo
    const int gl = first - 1 + tid;
    const bool owns = tid >= 1 && gl < n;
    const int i = gl < 0 ? n - 1 : gl;
    float e = 0.f, hi = 0.f, hip = 0.f;
    if (owns || tid == 0) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        hi = h[i];
        hip = h[ip];
        float v, dvdy;
        edge(i, ip, hi, hip, v, dvdy);
        const float dh = (hip - hi) * inv_dx;
        const float s2 = dh * dh;
        const float root = sqrtf(1.0f + s2);
        const float le = arc ? lt * (s2 / (root + 1.0f)) : 0.5f * lt * s2;
        if (owns) e = dxf * (le + v - f * hi);
        s_dvdy[tid] = dvdy;
        s_h[tid] = hi;
        s_slope[tid] = dh / root;
    }
    __syncthreads();
    if (owns && a.grad) {
        float lf;
        if (!arc) lf = -lt * ((s_h[tid - 1] + hip) - 2.0f * hi) * (inv_dx * inv_dx);
        else lf = lt * (s_slope[tid - 1] - s_slope[tid]) * inv_dx;
        static_cast<float *>(a.grad)[L.off + i] =
            dxf * (lf + 0.5f * (s_dvdy[tid - 1] + s_dvdy[tid]) - f);
    }
    if (a.energy) {
        const KSum k = block_ksum(KSum{e, 0.f}, red);
        if (tid == 0) static_cast<KSum *>(a.partial)[tile] = k;
    }
}

template <bool TRIAL, bool GRAD, bool ENERGY>
void launch64(hipStream_t s, const qew_host::EvalArgs &a, const LineDesc *one,
              const NoiseDesc *one_noise, int table_dtype, int grid) {
    LineDesc l{};
    NoiseDesc nd{};
    if (one) {
        l = *one;
        nd = *one_noise;
    }
    const int single = one ? 1 : 0;
    if (table_dtype == 1)
        hipLaunchKernelGGL((eval64_kernel<float, TRIAL, GRAD, ENERGY>), dim3(grid), dim3(kTile), 0, s,
                           a, l, nd, single);
    else
        hipLaunchKernelGGL((eval64_kernel<double, TRIAL, GRAD, ENERGY>), dim3(grid), dim3(kTile), 0,
                           s, a, l, nd, single);
}

__global__ void sample_kernel(NoiseDesc nd, int m, const double *x, const double *y, double *v,
                              double *dvdx, double *dvdy) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double a, b, c;
    if (nd.dtype == 1) sample_full64<float>(nd, x[k], y[k], a, b, c);
    else sample_full64<double>(nd, x[k], y[k], a, b, c);
    v[k] = a;
    dvdx[k] = b;
    dvdy[k] = c;
}

// DeviceNoise::sample_value (what = 0) or sample_dy (what = 1) at m points
// (src/filtered_noise.h:138-176): the value-only and dy-only samplers the
// objective and the gradient use.
__global__ void sample_part_kernel(NoiseDesc nd, int what, int m, const double *x,
                                   const double *y, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double r;
    if (nd.dtype == 1) r = what ? sample_dy64<float>(nd, x[k], y[k]) : sample_v64<float>(nd, x[k], y[k]);
    else r = what ? sample_dy64<double>(nd, x[k], y[k]) : sample_v64<double>(nd, x[k], y[k]);
    out[k] = r;
}

}  // namespace

namespace qew_host {

void launch_eval_f64(hipStream_t s, const EvalArgs &a, const LineDesc *one_line,
                     const NoiseDesc *one_noise) {
    // the table dtype of a batch must be uniform for the FP64 path; the batch
    // records it in noise 0 (checked at batch creation)
    const int dtype = one_noise ? one_noise->dtype : 0;
    const bool trial = a.d != nullptr, grad = a.grad != nullptr, energy = a.energy != nullptr;
    const int grid = a.ntiles;
    if (trial) {
        launch64<true, false, true>(s, a, one_line, one_noise, dtype, grid);
    } else if (grad && energy) {
        launch64<false, true, true>(s, a, one_line, one_noise, dtype, grid);
    } else if (grad) {
        launch64<false, true, false>(s, a, one_line, one_noise, dtype, grid);
    } else {
        launch64<false, false, true>(s, a, one_line, one_noise, dtype, grid);
    }
    QEW_HIP(hipGetLastError());
    if (energy) finish_energy_f64(s, a, one_line, 1);
}

void finish_energy_f64(hipStream_t s, const EvalArgs &a, const LineDesc *one_line, int nlines) {
    LineDesc l{};
    if (one_line) {
        l = *one_line;
        l.ntiles = a.ntiles;
    }
    hipLaunchKernelGGL(line_energy64, dim3(one_line ? 1 : nlines), dim3(256), 0, s, a.lines, l,
                       one_line ? 1 : 0, static_cast<const double *>(a.partial), a.energy);
    QEW_HIP(hipGetLastError());
}

void launch_eval_f32(hipStream_t s, const EvalArgs &a, int nlines) {
    hipLaunchKernelGGL(eval32_kernel, dim3(a.ntiles), dim3(kTile), 0, s, a);
    QEW_HIP(hipGetLastError());
    if (a.energy) {
        hipLaunchKernelGGL(line_energy32, dim3(nlines), dim3(256), 0, s, a.lines, a.anchor,
                           static_cast<const KSum *>(a.partial), a.energy);
        QEW_HIP(hipGetLastError());
    }
}

void ctx_reserve_eval(qew_ctx ctx, int ntiles, int nlines) {
    ctx->partial.ensure(sizeof(double) * (size_t)std::max(ntiles, 1));
    ctx->result.ensure(sizeof(double) * (size_t)std::max(nlines, 64));
    ctx->pinned.ensure(sizeof(double) * 64);
}

LineDesc single_line(int n, const qew_params *p) {
    LineDesc l{};
    l.off = 0;
    l.n = n;
    l.noise = 0;
    l.model = p->model;
    l.tile0 = 0;
    l.ntiles = (n + kTileSites - 1) / kTileSites;
    l.L = p->physical_size;
    l.lt = p->line_tension;
    l.f = p->driving_force;
    return l;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_scratch_init(int device) {
    QEW_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *ptr = nullptr;
    QEW_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    QEW_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
}

// The scratch address set up for `device`, or null.
void *noalloc_scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace qew_host

using namespace qew_host;

namespace {

void check_line(qew_ctx ctx, qew_noise noise, const qew_params *p, int n) {
    ctx_check(ctx);
    QEW_REQUIRE(noise != nullptr, "null noise");
    QEW_REQUIRE(p != nullptr, "null params");
    QEW_REQUIRE(n >= 1, "line length must be >= 1");
    QEW_REQUIRE(p->model == 0 || p->model == 1, "model must be QEW_LINEAR or QEW_ARCLENGTH");
    require_wide_table(noise->desc);
}

double single_eval(qew_ctx ctx, qew_noise noise, const qew_params *p, int n, const double *h,
                   const double *d, double alpha, double *grad, bool want_energy) {
    use_device(ctx);
    const LineDesc l = single_line(n, p);
    ctx_reserve_eval(ctx, l.ntiles, 1);
    EvalArgs a{};
    a.ntiles = l.ntiles;
    a.h = h;
    a.d = d;
    a.alpha = alpha;
    a.grad = grad;
    a.energy = want_energy ? ctx->result.as<double>() : nullptr;
    a.partial = ctx->partial.ptr;
    launch_eval_f64(ctx->stream, a, &l, &noise->desc);
    if (!want_energy) return 0.0;
    QEW_HIP(hipMemcpyAsync(ctx->pinned.ptr, ctx->result.ptr, sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    QEW_HIP(hipStreamSynchronize(ctx->stream));
    return *ctx->pinned.as<double>();
}

}  // namespace

extern "C" {

qew_status qew_objective(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                         const double *h, double *energy) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(h && energy, "null pointer");
        *energy = single_eval(ctx, noise, p, n, h, nullptr, 0.0, nullptr, true);
    });
}

qew_status qew_objective_trial(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                               const double *x0, const double *d, double alpha,
                               double *energy) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(x0 && d && energy, "null pointer");
        *energy = single_eval(ctx, noise, p, n, x0, d, alpha, nullptr, true);
    });
}

qew_status qew_gradient(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                        const double *h, double *grad) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(h && grad, "null pointer");
        single_eval(ctx, noise, p, n, h, nullptr, 0.0, grad, false);
    });
}

qew_status qew_gradient_edges(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                              const double *h, double *grad, double *edge_dvdy,
                              double *edge_slope) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(h && grad && edge_dvdy, "null pointer");
        use_device(ctx);
        const LineDesc l = single_line(n, p);
        ctx_reserve_eval(ctx, l.ntiles, 1);
        EvalArgs a{};
        a.ntiles = l.ntiles;
        a.h = h;
        a.grad = grad;
        a.partial = ctx->partial.ptr;
        a.edge_dvdy = edge_dvdy;
        a.edge_slope = edge_slope;
        launch_eval_f64(ctx->stream, a, &l, &noise->desc);
    });
}

qew_status qew_eval(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                    const double *h, double *grad, double *energy) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(h && grad, "null pointer");
        const double e = single_eval(ctx, noise, p, n, h, nullptr, 0.0, grad, energy != nullptr);
        if (energy) *energy = e;
    });
}

qew_status qew_eval_noalloc(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                            const double *h, double *grad, double *energy) {
    return guard([&] {
        check_line(ctx, noise, p, n);
        QEW_REQUIRE(h && grad, "null pointer");
        const LineDesc l = single_line(n, p);
        QEW_REQUIRE(l.ntiles <= kNoAllocMaxTiles,
                    "qew_eval_noalloc: a line of " + std::to_string(n) + " sites needs " +
                        std::to_string(l.ntiles) + " tiles of scratch; the maximum is " +
                        std::to_string(kNoAllocMaxTiles) + " tiles (" +
                        std::to_string(kNoAllocMaxTiles * kTileSites) + " sites)");
        void *scr = noalloc_scratch_ptr(ctx->device);
        QEW_REQUIRE(scr != nullptr, "the device's scratch address was not set up by qew_ctx_create");
        use_device(ctx);
        // every tile writes its partial before line_energy64 (the next launch
        // on the same stream) reads it: nothing needs clearing
        EvalArgs a{};
        a.ntiles = l.ntiles;
        a.h = h;
        a.grad = grad;
        a.energy = energy;
        a.partial = scr;
        launch_eval_f64(ctx->stream, a, &l, &noise->desc);
    });
}

qew_status qew_batch_eval(qew_batch b, int dtype, const void *h, const double *anchor,
                          void *grad, double *energy) {
    return guard([&] {
        QEW_REQUIRE(b != nullptr, "null batch");
        QEW_REQUIRE(h != nullptr, "null line buffer");
        QEW_REQUIRE(grad || energy, "nothing to compute");
        QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32, "dtype must be QEW_F64 or QEW_F32");
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        EvalArgs a{};
        a.lines = b->d_lines.as<LineDesc>();
        a.noises = b->d_noises.as<NoiseDesc>();
        a.tile_line = b->d_tile_line.as<int>();
        a.ntiles = b->ntiles;
        a.h = h;
        a.anchor = anchor;
        a.grad = grad;
        a.energy = energy;
        a.partial = b->d_partial.ptr;
        if (dtype == QEW_F32) {
            require_f32_tables(b->lines, b->noises, "QEW_F32 evaluation");
            launch_eval_f32(ctx->stream, a, b->nlines);
        } else {
            QEW_REQUIRE(!anchor, "anchors apply to QEW_F32 line state only");
            require_wide_tables(b->noises);
            NoiseDesc dtype_probe{};
            dtype_probe.dtype = b->noises.empty() ? 0 : b->noises[0].dtype;
            for (const auto &nd : b->noises)
                QEW_REQUIRE(nd.kind != TABLE || nd.dtype == dtype_probe.dtype,
                            "all tables of a batch must share one storage type");
            // batch mode: descriptors come from device arrays; pass the dtype only
            const bool trial = false;
            (void)trial;
            if (dtype_probe.dtype == 1) {
                if (grad && energy) launch64<false, true, true>(ctx->stream, a, nullptr, nullptr, 1, a.ntiles);
                else if (grad) launch64<false, true, false>(ctx->stream, a, nullptr, nullptr, 1, a.ntiles);
                else launch64<false, false, true>(ctx->stream, a, nullptr, nullptr, 1, a.ntiles);
            } else {
                if (grad && energy) launch64<false, true, true>(ctx->stream, a, nullptr, nullptr, 0, a.ntiles);
                else if (grad) launch64<false, true, false>(ctx->stream, a, nullptr, nullptr, 0, a.ntiles);
                else launch64<false, false, true>(ctx->stream, a, nullptr, nullptr, 0, a.ntiles);
            }
            QEW_HIP(hipGetLastError());
            if (energy) finish_energy_f64(ctx->stream, a, nullptr, b->nlines);
        }
    });
}

qew_status qew_noise_sample(qew_ctx ctx, qew_noise noise, int m, const double *x,
                            const double *y, double *v, double *dvdx, double *dvdy) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise != nullptr, "null noise");
        QEW_REQUIRE(m >= 0, "negative point count");
        if (m == 0) return;
        QEW_REQUIRE(x && y && v && dvdx && dvdy, "null pointer");
        require_wide_table(noise->desc);
        use_device(ctx);
        hipLaunchKernelGGL(sample_kernel, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                           noise->desc, m, x, y, v, dvdx, dvdy);
        QEW_HIP(hipGetLastError());
    });
}

static qew_status sample_part(qew_ctx ctx, qew_noise noise, int what, int m, const double *x,
                              const double *y, double *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise != nullptr, "null noise");
        QEW_REQUIRE(m >= 0, "negative point count");
        if (m == 0) return;
        QEW_REQUIRE(x && y && out, "null pointer");
        require_wide_table(noise->desc);
        use_device(ctx);
        hipLaunchKernelGGL(sample_part_kernel, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                           noise->desc, what, m, x, y, out);
        QEW_HIP(hipGetLastError());
    });
}

qew_status qew_noise_sample_value(qew_ctx ctx, qew_noise noise, int m, const double *x,
                                  const double *y, double *v) {
    return sample_part(ctx, noise, 0, m, x, y, v);
}

qew_status qew_noise_sample_dy(qew_ctx ctx, qew_noise noise, int m, const double *x,
                               const double *y, double *dvdy) {
    return sample_part(ctx, noise, 1, m, x, y, dvdy);
}

}  // extern "C"
