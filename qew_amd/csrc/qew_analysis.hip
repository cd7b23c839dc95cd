// Roughness analysis of a stack of periodic line profiles on the device: the
// height-difference autocorrelation, the power spectral density (both averaged
// over the stack) and the squared width of every profile. The reference leaves
// these to a Python module outside its checkout, called by its static
// post-processing (tools/postprocess_static.py:20-47, which writes `ell`,
// `acf`, `q`, `psd` beside each profile); the conventions are those of the
// host implementation in qew_amd/analysis.py:
//   W^2 = <(h - <h>)^2>,
//   C(q_k) = (L / n^2) |sum_x h(x) e^{-i q_k x}|^2,   q_k = 2 pi k / L,
//   A(r dx) = R(0) - R(r),  R(r) = (1/n) sum_x h~(x) h~(x + r),  h~ = h - <h>.
// One workgroup per profile centres it and forms W^2 (fixed-order workgroup
// sums); the batched global Stockham FFT transforms every centred profile; one
// pass averages |Z_k|^2 over the profiles in profile order; one more
// (unnormalised, inverse) FFT of that average gives n^2 R(r), because the
// stack average of R is the inverse transform of the stack-averaged power.
#include <algorithm>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kRB = 256;

__global__ __launch_bounds__(kRB) void centre_kernel(const double *h, int n, double2 *z, double *w2) {
    __shared__ double red[kRB / kWave];
    const int p = blockIdx.x;
    const double *hp = h + (size_t)p * n;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += kRB) s += hp[i];
    const double mean = block_sum(s, red) / (double)n;
    double v = 0;
    double2 *zp = z + (size_t)p * n;
    for (int i = threadIdx.x; i < n; i += kRB) {
        const double d = hp[i] - mean;
        zp[i] = make_double2(d, 0.0);
        v += d * d;
    }
    const double var = block_sum(v, red) / (double)n;
    if (w2 && threadIdx.x == 0) w2[p] = var;
}

// S_k = (1/m) sum_p |Z_pk|^2 (profiles in order) into spec[k] (imaginary part
// zero); C(q_k) = S_k L / n^2 for k = 1 .. n/2.
__global__ void power_kernel(const double2 *z, int m, int n, double L, double2 *spec, double *psd) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s = 0;
    for (int p = 0; p < m; ++p) {
        const double2 v = z[(size_t)p * n + k];
        s += v.x * v.x + v.y * v.y;
    }
    s /= (double)m;
    spec[k] = make_double2(s, 0.0);
    if (psd && k >= 1 && k <= n / 2) psd[k - 1] = s * (L / ((double)n * (double)n));
}

// A(r dx) = R(0) - R(r) for r = 1 .. n/2, with n^2 R(r) = Re(rn[r]).
__global__ void acf_kernel(const double2 *rn, int n, double *acf) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (r > n / 2) return;
    const double inv = 1.0 / ((double)n * (double)n);
    acf[r - 1] = rn[0].x * inv - rn[r].x * inv;
}

}  // namespace

using namespace qew_host;

extern "C" qew_status qew_roughness(qew_ctx ctx, int m, int n, double L, const double *h,
                                    double *acf, double *psd, double *w2) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(h, "null pointer");
        QEW_REQUIRE(m >= 1 && n >= 2, "need m >= 1 profiles of n >= 2 heights");
        QEW_REQUIRE(L > 0, "box length must be positive");
        use_device(ctx);
        hipStream_t s = ctx->stream;
        // workspace: z (m n), FFT scratch (m n), averaged spectrum (n) and its
        // FFT scratch (n), all FP64 complex
        const size_t mn = (size_t)m * (size_t)n;
        ctx->work.ensure(sizeof(double2) * (2 * mn + 2 * (size_t)n));
        double2 *z = ctx->work.as<double2>();
        double2 *tmp = z + mn;
        double2 *spec = tmp + mn;
        double2 *stmp = spec + n;
        hipLaunchKernelGGL(centre_kernel, dim3(m), dim3(kRB), 0, s, h, n, z, w2);
        QEW_HIP(hipGetLastError());
        if (!acf && !psd) return;
        fft_c2c(s, z, tmp, n, m, 1, n, -1);
        hipLaunchKernelGGL(power_kernel, dim3((n + 255) / 256), dim3(256), 0, s, z, m, n, L, spec,
                           psd);
        QEW_HIP(hipGetLastError());
        if (acf) {
            fft_c2c(s, spec, stmp, n, 1, 1, n, +1);
            hipLaunchKernelGGL(acf_kernel, dim3((n / 2 + 255) / 256), dim3(256), 0, s, spec, n,
                               acf);
            QEW_HIP(hipGetLastError());
        }
    });
}
