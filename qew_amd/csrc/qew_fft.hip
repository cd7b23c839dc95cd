// Global-memory Stockham autosort FFT, mixed radix (4, 2, 3, 5, and a direct
// DFT for any remaining prime factor up to 61), FP64, batched and strided.
// Used for the noise setup's 2-D transforms (src/filtered_noise.cpp:58,108,122)
// and the host-driven L-BFGS preconditioner (src/lbfgs.h:173-179). Each stage
// reads x and writes y (ping-pong), one launch per radix; twiddles come from
// sincospi on exact integer ratios. Work items are mapped so consecutive
// threads touch consecutive addresses: along the transform for unit-stride
// rows, across the batch for strided columns.
#include "qew_internal.hpp"

namespace {

constexpr int kMaxGenericRadix = 61;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 twiddle(int64_t num, int64_t den, int sign) {
    // exp(sign * 2 pi i num / den)
    double s, c;
    sincospi(2.0 * (double)num / (double)den, &s, &c);
    return make_double2(c, sign * s);
}

template <int R>
__device__ __forceinline__ void dft_small(double2 v[R], int sign) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const double2 a = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        const double2 b = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        const double2 c = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        const double2 e = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        // (sign * i) * e
        const double2 je = (sign < 0) ? make_double2(e.y, -e.x) : make_double2(-e.y, e.x);
        v[0] = make_double2(a.x + c.x, a.y + c.y);
        v[1] = make_double2(b.x + je.x, b.y + je.y);
        v[2] = make_double2(a.x - c.x, a.y - c.y);
        v[3] = make_double2(b.x - je.x, b.y - je.y);
    } else {
        double2 out[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            double2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const double2 w = twiddle((int64_t)((q * s) % R), R, sign);
                acc.x += v[q].x * w.x - v[q].y * w.y;
                acc.y += v[q].x * w.y + v[q].y * w.x;
            }
            out[s] = acc;
        }
#pragma unroll
        for (int s = 0; s < R; ++s) v[s] = out[s];
    }
}

struct StageArgs {
    const double2 *x;
    double2 *y;
    int n, Ns, batch, R;
    int64_t stride, bstride;
    int sign;
    int along;  // 1: consecutive threads walk the transform, 0: the batch
};

__device__ __forceinline__ bool stage_index(const StageArgs &a, int R, int64_t t, int &j, int &b) {
    const int64_t m = a.n / R;
    if (t >= m * (int64_t)a.batch) return false;
    if (a.along) {
        j = (int)(t % m);
        b = (int)(t / m);
    } else {
        b = (int)(t % a.batch);
        j = (int)(t / a.batch);
    }
    return true;
}

template <int R>
__global__ __launch_bounds__(256) void stockham_stage(StageArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int j, b;
    if (!stage_index(a, R, t, j, b)) return;
    const int m = a.n / R;
    const int k = j % a.Ns;
    const double2 *x = a.x + (int64_t)b * a.bstride;
    double2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = x[(int64_t)(j + q * m) * a.stride];
    if (a.Ns > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[q] = cmul(v[q], twiddle((int64_t)q * k, (int64_t)a.Ns * R, a.sign));
    }
    dft_small<R>(v, a.sign);
    double2 *y = a.y + (int64_t)b * a.bstride;
    const int64_t idxD = (int64_t)(j / a.Ns) * a.Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) y[(idxD + (int64_t)q * a.Ns) * a.stride] = v[q];
}

__global__ __launch_bounds__(256) void stockham_stage_generic(StageArgs a) {
    const int R = a.R;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int j, b;
    if (!stage_index(a, R, t, j, b)) return;
    const int m = a.n / R;
    const int k = j % a.Ns;
    const double2 *x = a.x + (int64_t)b * a.bstride;
    double2 v[kMaxGenericRadix];
    for (int q = 0; q < R; ++q) {
        double2 u = x[(int64_t)(j + q * m) * a.stride];
        if (a.Ns > 1 && q > 0) u = cmul(u, twiddle((int64_t)q * k, (int64_t)a.Ns * R, a.sign));
        v[q] = u;
    }
    double2 *y = a.y + (int64_t)b * a.bstride;
    const int64_t idxD = (int64_t)(j / a.Ns) * a.Ns * R + k;
    for (int s = 0; s < R; ++s) {
        double2 acc = v[0];
        for (int q = 1; q < R; ++q) {
            const double2 w = twiddle((int64_t)((q * s) % R), R, a.sign);
            acc.x += v[q].x * w.x - v[q].y * w.y;
            acc.y += v[q].x * w.y + v[q].y * w.x;
        }
        y[(idxD + (int64_t)s * a.Ns) * a.stride] = acc;
    }
}

std::vector<int> factor(int n) {
    std::vector<int> r;
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    while (n % 3 == 0) { r.push_back(3); n /= 3; }
    while (n % 5 == 0) { r.push_back(5); n /= 5; }
    for (int p = 7; n > 1; p += 2) {
        while (n % p == 0) {
            if (p > kMaxGenericRadix)
                throw qew_host::Error(QEW_ERR_UNSUPPORTED,
                                      "FFT length has a prime factor above 61: " + std::to_string(p));
            r.push_back(p);
            n /= p;
        }
    }
    return r;
}

}  // namespace

namespace qew_host {

void fft_c2c(hipStream_t s, double2 *data, double2 *tmp, int n, int batch, int64_t stride,
             int64_t bstride, int sign) {
    if (n <= 1 || batch <= 0) return;
    const std::vector<int> rs = factor(n);
    double2 *src = data, *dst = tmp;
    int Ns = 1;
    for (int R : rs) {
        StageArgs a{src, dst, n, Ns, batch, R, stride, bstride, sign, stride == 1 ? 1 : 0};
        const int64_t work = (int64_t)(n / R) * batch;
        const dim3 grid((unsigned)((work + 255) / 256));
        switch (R) {
        case 2: hipLaunchKernelGGL(stockham_stage<2>, grid, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(stockham_stage<3>, grid, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL(stockham_stage<4>, grid, dim3(256), 0, s, a); break;
        case 5: hipLaunchKernelGGL(stockham_stage<5>, grid, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL(stockham_stage_generic, grid, dim3(256), 0, s, a); break;
        }
        QEW_HIP(hipGetLastError());
        std::swap(src, dst);
        Ns *= R;
    }
    // an odd number of stages leaves the result in tmp
    if (src != data) {
        if (stride == 1 && bstride == n) {
            QEW_HIP(hipMemcpyAsync(data, src, sizeof(double2) * (size_t)n * batch,
                                   hipMemcpyDeviceToDevice, s));
        } else {
            // strided layout: copy the whole spanned extent
            const int64_t span = (int64_t)(n - 1) * stride + (int64_t)(batch - 1) * bstride + 1;
            QEW_HIP(hipMemcpyAsync(data, src, sizeof(double2) * (size_t)span,
                                   hipMemcpyDeviceToDevice, s));
        }
    }
}

}  // namespace qew_host
